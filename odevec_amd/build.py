"""Build libodevec.so for gfx950.

1. Run the preprocessor for the built-in network variants (and any extra
   ``NETWORK/ORDERING/PACKAGING`` ids) into ``csrc/generated``.
2. Compile each instantiation unit and the C ABI with hipcc (one process per
   unit, in parallel) and link ``odevec_amd/libodevec.so``.
3. Compile the Fortran module (``fortran/odevec_main.F90``,
   ``fortran/odevec_commons.F90``) and the ROBER driver program against it
   (``fortran/_build/rober_driver``), if a Fortran compiler is present.

    python -m odevec_amd.build [--extra PRIMORDIAL/INVERT/CSC] [-j 8]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from typing import List

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "generated")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libodevec.so")

BUILTIN_VARIANTS = [
    "ROBER/NONE/DENSE",
    "ROBER/NONE/CSC",
    "OREGO/NONE/DENSE",
    "OREGO/NONE/CSC",
    "PRIMORDIAL/CMK/CSC",
    "PRIMORDIAL/NONE/DENSE",
    "PRIMORDIAL/FF/CSC",
    "SYNTH64/FF/CSC",     # 64 species: cooperative kernels only
    "SYNTH40/FF/CSC",     # 40 species: cooperative kernels, 24 padding lanes per group
    "ROBER/NONE/DENSE/LU3",        # LUmethod 3: Jacobian by forward differences
    "PRIMORDIAL/NONE/DENSE/LU3",
    "ROBER/NONE/DENSE/LU2",        # LUmethod 2: L and U in closed form
    "OREGO/NONE/DENSE/LU2",
    "KROME/FF/CSC",                # the KROME setup file KROME_SETUPFILE
]

# the network that KROME/... variants are generated from: a KROME-style setup
# file (Python syntax, read as pre_odevec.py:97-105 reads one)
KROME_SETUPFILE = os.path.join(HERE, "pre_odevec", "krome_ism.py")

# bundles: one batch kernel (odv_k_batch) over several (network, precision)
# entries; every network listed must also be a variant above. The kernel's
# register allocation covers all its entries, so a bundle with one
# precision per network spills less than one with both: the autotuner picks
# the bundle (the register/LDS split) and the block size per batch.
_MIX3 = ["ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC"]
BUNDLES = {
    "MIX3": [(n, p) for n in _MIX3 for p in (0, 1)],                          # every split
    "MIX3_FP64": [(n, 0) for n in _MIX3],                                     # all FP64
    "MIX3_MIXED": [(n, 1 if n.startswith("PRIMORDIAL") else 0) for n in _MIX3],  # FP32 LU for PRIMORDIAL
}

# the reproducible build: no contraction of a*b+c into FMA, so device results
# equal the oracle's bit for bit
REPRO_FLAGS = ["-ffp-contract=off", "-DODV_REPRO=1"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions (hipcc's host pass warns about the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def generate(variants: List[str], krome_setupfile: str = None) -> List[str]:
    from .pre_odevec import cli
    bases = []
    for v in variants:
        name, ordering, packaging, lu = _parts(v)
        ordering = "None" if ordering.upper() == "NONE" else ordering
        kf = (krome_setupfile or KROME_SETUPFILE) if name.upper() == "KROME" else None
        cli.generate(name, ordering, packaging, GEN, krome_setupfile=kf, lu_method=lu)
        bases.append("%s_%s_%s%s" % (name, ordering.upper(), packaging, "" if lu == 1 else "_LU%d" % lu))
    return bases


def _parts(variant: str):
    """NAME/ORDERING/PACKAGING[/LU2|/LU3] -> (name, ordering, packaging, LUmethod)."""
    p = variant.split("/")
    if len(p) == 4 and p[3] in ("LU2", "LU3"):
        return p[0], p[1], p[2], int(p[3][2])
    if len(p) != 3:
        raise ValueError("variant %r: expected NAME/ORDERING/PACKAGING[/LU2|/LU3]" % variant)
    return p[0], p[1], p[2], 1


def _struct(variant: str) -> str:
    name, ordering, packaging, lu = _parts(variant)
    return "odv_net_%s_%s_%s%s" % (name, ordering.upper(), packaging, "" if lu == 1 else "_LU%d" % lu)


def generate_bundle(name: str, entries) -> str:
    """Instantiation unit of bundle `name`: its (network, precision) entries
    (precision 0 FP64, 1 mixed), registered under the bundle's name."""
    nets = sorted(set(n for n, _ in entries), key=[n for n, _ in entries].index)
    lines = ["// Generated by odevec_amd.build -- do not edit.",
             "// bundle %s: %s" % (name, ", ".join("%s %s" % (n, "mixed" if p else "FP64") for n, p in entries)),
             '#include "odevec_registry.h"']
    for n in nets:
        lines.append('#include "generated/net_%s.h"' % _struct(n)[len("odv_net_"):])
    ent = ",\n    ".join("odv::Entry<%s, %s, %s>" % (_struct(n), "float" if p else "double", "true" if p else "false")
                          for n, p in entries)
    lines += ["namespace {",
              "using Bundle_%s = odv::Bundle<\n    %s>;" % (name, ent),
              "const char* const nets_%s[] = {%s};" % (name, ", ".join('"%s"' % n for n, _ in entries)),
              "const int precs_%s[] = {%s};" % (name, ", ".join(str(p) for _, p in entries)),
              "const int neq_%s[] = {%s};" % (name, ", ".join("%s::NEQ" % _struct(n) for n, _ in entries)),
              "const int npar_%s[] = {%s};" % (name, ", ".join("%s::NPARAM" % _struct(n) for n, _ in entries)),
              "void (*const defs_%s[])(double*) = {%s};" % (
                  name, ", ".join("&%s::param_defaults" % _struct(n) for n, _ in entries)),
              "}  // namespace",
              "ODV_REGISTER_BUNDLE(%s, Bundle_%s, nets_%s, precs_%s, neq_%s, npar_%s, defs_%s)" % ((name,) * 7),
              ""]
    path = os.path.join(GEN, "inst_bundle_%s.hip" % name)
    text = "\n".join(lines)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def _compile(src: str, obj: str, extra: List[str]) -> str:
    cmd = [hipcc()] + HIPCC_FLAGS + extra + ["-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr[-6000:]))
    return obj


def build(variants: List[str] = None, jobs: int = 8, verbose: bool = True, repro: bool = True,
          bundles: bool = True, krome_setupfile: str = None) -> str:
    variants = list(BUILTIN_VARIANTS if variants is None else variants)
    bases = generate(variants, krome_setupfile)
    os.makedirs(BUILD, exist_ok=True)
    units = [(os.path.join(CSRC, "odevec_capi.hip"), os.path.join(BUILD, "odevec_capi.o"), [])]
    for b in bases:
        src = os.path.join(GEN, "inst_%s.hip" % b)
        units.append((src, os.path.join(BUILD, "inst_%s.o" % b), []))
        if repro:
            units.append((src, os.path.join(BUILD, "inst_%s_repro.o" % b), REPRO_FLAGS))
    for name, entries in (BUNDLES.items() if bundles else []):
        if not all(n in variants for n, _ in entries):
            continue
        src = generate_bundle(name, entries)
        units.append((src, os.path.join(BUILD, "bundle_%s.o" % name), []))
        if repro:
            units.append((src, os.path.join(BUILD, "bundle_%s_repro.o" % name), REPRO_FLAGS))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, s, o, f) for s, o, f in units]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built %s (%d network variants%s)" % (LIB, len(bases), ", each also +REPRO" if repro else ""))
    if "ROBER/NONE/DENSE" in variants and repro:
        build_fortran(verbose=verbose)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "_build")
FORTRAN_DRIVER = os.path.join(FORTRAN_BUILD, "rober_driver")


def flang() -> str:
    for c in (os.environ.get("FLANG"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return ""


def build_fortran(network: str = "ROBER/NONE/DENSE+REPRO", nvector: int = 32, neq: int = 3,
                  verbose: bool = True) -> str:
    """The Fortran module (odevec_main, odevec_commons) and the ROBER driver
    program linked against libodevec.so (the driver uses the reproducible
    build, so its output can be checked bit for bit); skipped if no Fortran
    compiler."""
    fc = flang()
    if not fc:
        if verbose:
            print("no Fortran compiler: Fortran module not built")
        return ""
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    defs = ["-DODEVEC_NETWORK=\"%s\"" % network, "-DODEVEC_NVECTOR=%d" % nvector, "-DODEVEC_NEQ=%d" % neq,
            "-DODEVEC_PRECISION=0", "-DODEVEC_MODE=1"]
    objs = []
    for src in ("odevec_main.F90", "odevec_commons.F90", "rober_driver.f90"):
        obj = os.path.join(FORTRAN_BUILD, src.split(".")[0] + ".o")
        cmd = [fc, "-cpp", "-O2", "-J", FORTRAN_BUILD] + defs + ["-c", os.path.join(FORTRAN, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("Fortran compile failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
        objs.append(obj)
    cmd = [fc, "-o", FORTRAN_DRIVER] + objs + ["-L", HERE, "-lodevec", "-Wl,-rpath,$ORIGIN/../.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    if verbose:
        print("built %s" % FORTRAN_DRIVER)
    return FORTRAN_DRIVER


def main(argv=None) -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--extra", action="append", default=[], help="NETWORK/ORDERING/PACKAGING to add")
    p.add_argument("-j", "--jobs", type=int, default=8)
    p.add_argument("--krome_setupfile", default=None,
                   help="KROME-style setup file for the KROME/... variants (default %s)" % KROME_SETUPFILE)
    a = p.parse_args(argv)
    build(BUILTIN_VARIANTS + a.extra, a.jobs, krome_setupfile=a.krome_setupfile)
    return 0


if __name__ == "__main__":
    sys.exit(main())
