"""Command line of the preprocessor, flag-compatible with the reference's
``pre_odevec.py`` (pre_odevec.py:383-448) where the flag still has a meaning.

    python -m odevec_amd.pre_odevec --nvector=32 --example=ROBER
    python -m odevec_amd.pre_odevec --nvector=64 --example=PRIMORDIAL --ordering=CMK --packaging=CSC

Instead of Fortran files it writes ``net_<NETWORK>_<ORDERING>_<PACKAGING>.h``
(HIP/C++ device code), the network IR ``<same>.json`` and the
instantiation unit ``inst_<same>.hip`` into ``--outdir``; rebuilding the
library (``python -m odevec_amd.build --extra <id>``) adds the network to it.

Flag notes:
* ``--nvector`` is accepted for compatibility; the cell count is a runtime
  size of the library, not a code-generation constant.
* ``--LUmethod`` 1 (symbolic Jacobian + generated LU), 2 (the L and U
  factors in closed form, pre_odevec.py:132-139, variant id
  ``NAME/ORDERING/DENSE/LU2``) and 3 (Jacobian by forward differences of the
  RHS, src/odevec.F90:469-506, then the generated LU; variant id
  ``NAME/NONE/DENSE/LU3``) are generated. 2 needs DENSE packaging: the
  reference's CSC form writes the factors in SymPy's column-list order, which
  is not the order its CSC solve reads them in. 3 needs DENSE packaging (the
  reference aborts for CSC, src/odevec.F90:509-523) and no ordering (the
  reference permutes the residual but not this Jacobian, src/odevec.F90:306-308,
  328-342, which is only consistent without one).
* ``--heatcool 1`` needs KROME's heating/cooling Fortran and is rejected.
* ``--equilibrium_H`` is validated exactly as the reference validates it
  (pre_odevec.py:455-460: DENSE, no ordering, LUmethod 3) and, as there,
  changes nothing else in the generated code.
* ``--dt_min`` is recorded in the IR as the network's default MinimalTimestep.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from typing import List, Optional

from . import networks, structure, emit


def generate(example: str, ordering: str, packaging: str, outdir: str,
             krome_setupfile: Optional[str] = None, dt_min: Optional[float] = None,
             quiet: bool = True, lu_method: int = 1) -> dict:
    net = networks.load(example, krome_setupfile)
    st = structure.analyse(net, ordering, packaging)
    if lu_method not in (1, 2, 3):
        raise ValueError("LUmethod %d is not generated (1: symbolic Jacobian, 2: closed-form LU, "
                         "3: finite differences)" % lu_method)
    if lu_method == 2 and st.packaging != "DENSE":
        raise ValueError("LUmethod 2 needs DENSE packaging")
    if lu_method == 3 and (st.packaging != "DENSE" or st.ordering != "None"):
        raise ValueError("LUmethod 3 needs DENSE packaging and no ordering")
    st.lu_method = lu_method
    os.makedirs(outdir, exist_ok=True)
    base = "%s_%s_%s%s" % (net.name, st.ordering.upper(), st.packaging, emit.lu_suffix(lu_method, "_"))
    hdr = emit.emit_header(st)
    ir = emit.ir_dict(st)
    if dt_min is not None:
        ir["dt_min"] = float(dt_min)
    files = {
        "net_%s.h" % base: hdr,
        "%s.json" % base: json.dumps(ir, indent=1) + "\n",
        "inst_%s.hip" % base: (
            "// Generated by odevec_amd.pre_odevec -- do not edit.\n"
            '#include "../odevec_registry.h"\n'
            '#include "net_%s.h"\n'
            "ODV_REGISTER(%s)\n" % (base, emit.struct_name(st))),
    }
    for name, text in files.items():
        path = os.path.join(outdir, name)
        old = None
        if os.path.exists(path):
            with open(path) as fh:
                old = fh.read()
        if old != text:
            with open(path, "w") as fh:
                fh.write(text)
    if not quiet:
        print("#  network %s: neq %d, nnz(J) %d, nnz(P) %d, nnz(LU) %d" % (
            net.name, st.neq, int(st.J.sum()), int(st.P.sum()), st.csc.nnz))
        print("#  permutation", [int(p) for p in st.perm])
        print("#  wrote", ", ".join(files))
    return ir


def main(argv: Optional[List[str]] = None) -> int:
    p = argparse.ArgumentParser(
        description="Preprocessor for the MI355X OdeVec integrator: emits HIP device code "
                    "for a network's RHS, Jacobian and sparse LU.")
    p.add_argument("--example", default="PRIMORDIAL", help="ROBER, OREGO, PRIMORDIAL, SYNTH64, SYNTH40 or KROME")
    p.add_argument("--krome_setupfile", default=None, help="KROME-style network file (Python syntax)")
    p.add_argument("--LUmethod", type=int, default=1)
    p.add_argument("--dt_min", type=float, default=float(__import__("numpy").float32(1e-10)))
    p.add_argument("--nvector", type=int, default=1, required=False)
    p.add_argument("--heatcool", type=int, default=0)
    p.add_argument("--equilibrium_H", type=int, default=0)
    p.add_argument("--ordering", default="None", help="None, CMK, INVERT or FF")
    p.add_argument("--packaging", default="DENSE", help="DENSE or CSC")
    p.add_argument("--outdir", default=os.path.join(os.path.dirname(__file__), "..", "csrc", "generated"))
    p.add_argument("--sparsity_structure", default="False")
    args = p.parse_args(argv)
    if args.krome_setupfile is not None:
        args.example = "KROME"
    if args.equilibrium_H == 1 and args.packaging != "DENSE":
        sys.exit("args.equilibrium_H is only allowed with DENSE packaging.")
    if args.equilibrium_H == 1 and args.ordering != "None":
        sys.exit("args.equilibrium_H is only allowed without ordering.")
    if args.equilibrium_H == 1 and args.LUmethod != 3:
        sys.exit("args.equilibrium_H is currently only usable with LUmethod==3.")
    if args.LUmethod not in (1, 2, 3):
        sys.exit("LUmethod %d is not generated; 1 (symbolic Jacobian), 2 (closed-form LU) and 3 "
                 "(finite differences) are." % args.LUmethod)
    if args.LUmethod == 2 and args.packaging != "DENSE":
        sys.exit("LUmethod 2 needs DENSE packaging.")
    if args.LUmethod == 3 and (args.packaging != "DENSE" or args.ordering not in ("None", "NONE")):
        sys.exit("LUmethod 3 needs DENSE packaging and no ordering.")
    if args.heatcool:
        sys.exit("--heatcool needs KROME's heating/cooling routines, which are not part of this preprocessor.")
    if args.packaging not in structure.PACKAGINGS:
        sys.exit("# Error: This format is not known.")
    ir = generate(args.example, args.ordering, args.packaging, args.outdir,
                  args.krome_setupfile, args.dt_min, quiet=False, lu_method=args.LUmethod)
    if args.sparsity_structure == "True":
        n = ir["neq"]
        lu = {tuple(e) for e in ir["lu_entries"]}
        print("#  sparsity structure of LU (ordered):")
        for i in range(n):
            print("   " + "".join("X" if (i, j) in lu else "." for j in range(n)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
