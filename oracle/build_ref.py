"""Recipe for oracle/_ref: the reference's own brick and shell leaf kernels,
compiled from their sources where they lie under the reference tree, plus the
test harnesses oracle/ref/harness_brick.F, harness_shell.F, harness_qeph.F,
harness_qbat.F, harness_nodal.F and harness_int7.F, linked into
oracle/_ref/libref.so.

Test infrastructure only.  Runs in the build container (amdflang from ROCm);
outputs go to oracle/_ref/ only.  Module dependencies are resolved by compiling
each missing module's defining file from the reference tree, recursively.

Flags follow the reference's LLVM-flang double-precision recipe
(engine/CMake_Compilers/cmake_linuxa64.txt:83-103): -DMYREAL8, fixed-form with
unlimited line length, no FMA contraction, no fast-math.
"""
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
REF = os.environ.get("ORAD_REFERENCE_ROOT", "/root/reference")

LEAF = [
    "engine/source/elements/solid/solide/srcoor3.F",
    "engine/source/elements/solid/solide/srepiso3.F",
    "engine/source/elements/solid/solide/sortho3.F",
    "engine/source/elements/solid/solide/srrota3.F",
    "engine/source/elements/solid/solide/sderi3.F",
    "engine/source/elements/solid/solide4/schkjabt3.F",
    "engine/source/elements/solid/solide/sdlen3.F",
    "engine/source/elements/solid/solide/slen.F",
    "engine/source/elements/solid/solide/sdefo3.F",
    "engine/source/elements/solid/solide/srho3.F",
    "engine/source/elements/solid/solide/srmall3.F",
    "engine/source/materials/mat/mat002/m2law.F",
    "engine/source/materials/mat/mat002/m2iter_imp.F",
    "engine/source/materials/mat_share/mqviscb.F",
    "engine/source/materials/mat_share/mstrain_rate.F",
    "engine/source/elements/solid/solide/shvis3.F",
    "engine/source/elements/solid/solide/sfint3.F",
    "engine/source/elements/solid/solide/scumu3p.F",
    "engine/source/elements/shell/coque/ccoor3.F",
    "engine/source/elements/shell/coque/cnvec3.F",
    "engine/source/elements/shell/coque/cortdir3.F",
    "engine/source/elements/shell/coque/cderi3.F",
    "engine/source/elements/shell/coque/ccoef3.F",
    "engine/source/elements/shell/coque/cdlen3.F",
    "engine/source/elements/shell/coque/cdefo3.F",
    "engine/source/elements/shell/coque/ccurv3.F",
    "engine/source/elements/shell/coque/cstra3.F",
    "engine/source/elements/shell/coque/chvis3.F",
    "engine/source/elements/shell/coque/cdt3.F",
    "engine/source/elements/shell/coque/cfint3.F",
    "engine/source/elements/shell/coque/cupdt3.F",
    "engine/source/elements/shell/coqini.F",
    "engine/source/materials/mat/mat002/sigeps02c.F",
    "engine/source/materials/mat/mat002/m2cplr.F",
    "engine/source/elements/shell/coquez/czcorc.F",
    "engine/source/elements/shell/coquez/czcorp5.F",
    "engine/source/elements/shell/coquez/a3invdp_v.F",
    "engine/source/elements/sh3n/coquedk/cdkcoor3.F",
    "engine/source/elements/sh3n/coquedk/cncoef3.F",
    "engine/source/elements/shell/coquez/czdef.F",
    "engine/source/elements/shell/coquez/czstra3.F",
    "engine/source/elements/sh3n/coquedk/cndt3.F",
    "engine/source/elements/shell/coquez/czfintce.F",
    "engine/source/elements/shell/coquez/czfintn.F",
    "engine/source/elements/shell/coquez/czproj.F",
    "engine/source/elements/shell/coque/cupdtn3.F",
    "engine/source/elements/shell/coqueba/cbacoor.F",
    "engine/source/elements/shell/coqueba/cbadef.F",
    "engine/source/elements/shell/coqueba/cbastra3.F",
    "engine/source/elements/shell/coqueba/cbaener.F",
    "engine/source/elements/shell/coqueba/cbavisc.F",
    "engine/source/elements/shell/coqueba/cbafori.F",
    "engine/source/elements/shell/coqueba/cbaproj.F",
    "engine/source/assembly/accele.F",
    "engine/source/constraints/general/bcs/bcs10.F",
    "engine/source/constraints/general/bcs/bcs1.F",
    "engine/source/constraints/general/impvel/fixvel.F",
    "engine/source/tools/curve/vinterdp.F",
    "engine/source/assembly/velocity.F",
    "engine/source/assembly/displacement.F",
    "engine/source/interfaces/int07/i7cor3.F",
    "engine/source/interfaces/int07/i7dst3.F",
    "engine/source/interfaces/int07/i7for3.F",
    "engine/source/interfaces/int07/i7ass3.F",
]


# C++ sources of the reference that the Fortran routines above call (FIXVEL's
# function-evaluation path reaches the hash map and the Python-function modules);
# compiled with g++ from where they lie, only standard headers needed.
CXX = [
    "common_source/tools/container/umap.cpp",
    "common_source/modules/cpp_python_funct.cpp",
]


def flags():
    inc = [
        f"-I{REF}/engine/share/includes", f"-I{REF}/engine/share/spe_inc",
        f"-I{REF}/engine/share/r8", f"-I{REF}/common_source/includes",
        f"-I{REF}/engine/share/message_module", f"-J{OUT}", f"-I{OUT}",
    ]
    return ["amdflang", "-c", "-fPIC", "-cpp", "-ffixed-line-length-none",
            "-DMYREAL8", "-DCPP_mach=CPP_p4linux964", "-O2",
            "-ffp-contract=off", "-fno-fast-math"] + inc


def module_index():
    idx = {}
    for root in ("common_source", "engine/share", "engine/source"):
        for f in glob.glob(f"{REF}/{root}/**/*.F*", recursive=True):
            try:
                txt = open(f, errors="ignore").read()
            except OSError:
                continue
            for m in re.findall(r"^\s*module\s+(\w+)\s*$", txt, re.I | re.M):
                idx.setdefault(m.lower(), f)
    return idx


def main():
    if not os.path.isdir(REF):
        print(f"reference tree {REF} absent: oracle/_ref not built")
        return 0
    os.makedirs(OUT, exist_ok=True)
    # Start clean: a stale .mod would hide its object from the link list.
    for f in glob.glob(os.path.join(OUT, "*.o")) + glob.glob(os.path.join(OUT, "*.mod")):
        os.remove(f)
    fl = flags()
    idx = module_index()
    objs, done = [], set()

    def comp(src):
        if src in done:
            return
        obj = os.path.join(OUT, re.sub(r"[^A-Za-z0-9_]", "_",
                                       os.path.relpath(src, REF)) + ".o")
        for _ in range(64):
            p = subprocess.run(fl + [src, "-o", obj], capture_output=True, text=True)
            if p.returncode == 0:
                done.add(src)
                objs.append(obj)
                return
            missing = re.findall(r"Cannot parse module file for module '(\w+)'", p.stderr)
            if not missing:
                raise RuntimeError(f"compile failed: {src}\n{p.stderr[:3000]}")
            for m in missing:
                if m.lower() not in idx:
                    raise RuntimeError(f"module {m} has no source in the reference tree")
                comp(idx[m.lower()])
        raise RuntimeError(f"dependency loop on {src}")

    for rel in LEAF:
        comp(os.path.join(REF, rel))
    comp(os.path.join(HERE, "ref", "harness_brick.F"))
    comp(os.path.join(HERE, "ref", "harness_shell.F"))
    comp(os.path.join(HERE, "ref", "harness_qeph.F"))
    comp(os.path.join(HERE, "ref", "harness_qbat.F"))
    comp(os.path.join(HERE, "ref", "harness_nodal.F"))
    comp(os.path.join(HERE, "ref", "harness_int7.F"))
    for rel in CXX:
        obj = os.path.join(OUT, re.sub(r"[^A-Za-z0-9_]", "_", rel) + ".o")
        subprocess.run(["g++", "-c", "-fPIC", "-O2", "-std=c++17", os.path.join(REF, rel),
                        "-o", obj], check=True)
        objs.append(obj)
    lib = os.path.join(OUT, "libref.so")
    # Routines referenced on branches the path does not take (error messages,
    # SPH, orthotropic frames) are compiled from their own reference sources too,
    # until the library has no unresolved Fortran symbol.
    # ANCMSG (message catalogue) and ARRET (abort, restart flush) are only reached
    # on error branches and pull in the whole I/O layer, including C++ sources that
    # need headers absent from the image; they and the routines of ERROR_PATH below
    # that no harness reaches are left weak, so a call would fault loudly rather
    # than run a substitute.
    subs = subroutine_index()
    for _ in range(32):
        subprocess.run(["amdflang", "-shared", "-o", lib] + objs + LIBS, check=True)
        und = subprocess.run(["nm", "-u", lib], capture_output=True, text=True).stdout.split()
        need = sorted({s[:-1] for s in und if re.fullmatch(r"[a-z][a-z0-9_]*_", s)} - {""})
        need = [s for s in need if s in subs and s not in ERROR_PATH]
        if not need:
            break
        for s in need:
            comp(subs[s])
    weak = []
    for s in ERROR_PATH:
        weak += ["--weaken-symbol", s + "_"]
    for s in C_ERROR_PATH:
        weak += ["--weaken-symbol", s]
    for o in objs:
        subprocess.run(["objcopy"] + weak + [o], check=True)
    subprocess.run(["amdflang", "-shared", "-o", lib] + objs + LIBS, check=True)
    und = subprocess.run(["nm", "-u", lib], capture_output=True, text=True).stdout.splitlines()
    left = [ln.split()[-1] for ln in und
            if ln.split()[0] == "U" and re.fullmatch(r"[a-z][a-z0-9_]*_", ln.split()[-1])]
    if left:
        raise RuntimeError(f"unresolved reference routines: {left}")
    print(f"built {lib} from {len(objs)} objects")
    return 0


LIBS = ["-lstdc++", "-ldl"]

ERROR_PATH = ("ancmsg", "arret",
              # Called only by CNCOEF3 (sh3n/coquedk/cncoef3.F:610-939), the
              # coefficient routine of the DKT/QBAT shells, which shares its source
              # file with CNCOEF3B; their own file (coqueba/cmatc3.F) pulls in the
              # implicit solver's initialisation and from there the whole engine.
              "gepm_lc", "cctoglob",
              # Restart-file readers and writers in C (common_source/tools/
              # input_output/write_routines.c, which needs zlib) referenced by the
              # contact modules' restart routines; no harness reads or writes a
              # restart file.
              "read_i_c", "write_i_c", "read_i_array_c", "write_i_array_c")


# C functions bound by OUTPUT_MOD's checksum module (common_source/output/checksum/
# checksum.cpp, which needs an md5 header absent from the image); only the output
# layer calls them, never the contact routines on the harness's path.
C_ERROR_PATH = ("new_file_checksum_list", "print_checksum_list")


def subroutine_index():
    idx = {}
    for root in ("engine/source", "common_source", "engine/share"):
        for f in glob.glob(f"{REF}/{root}/**/*.F*", recursive=True):
            try:
                txt = open(f, errors="ignore").read()
            except OSError:
                continue
            for m in re.findall(r"^\s+(?:recursive\s+)?subroutine\s+(\w+)", txt, re.I | re.M):
                idx.setdefault(m.lower(), f)
    return idx


if __name__ == "__main__":
    sys.exit(main())
