"""MI355X-native explicit element loop (bricks) behind an OpenRadioss-style C ABI."""
