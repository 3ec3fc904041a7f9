"""Copy the data fixtures that HONEI's own tests use for the ELL solve path
into tests/golden/honei_testdata/ (run once, in a container where the
reference tree is mounted).  Only data files are copied -- matrices in the
reference's ELL binary format and vectors in its EXP text format -- never
source code.

    python tests/golden/copy_reference_fixtures.py [/root/reference]
"""
import os
import shutil
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "honei/math/testdata")
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "honei_testdata")

FILES = (
    # CSR->ELL golden pair (clients/io_tools/mtxtoell.cc wrote the .ell)
    ["5pt_10x10.mtx", "5pt_10x10.ell"]
    # FEAST Q1 hierarchy + FEAT solution used by math/mg_TEST.cc:389-464
    + [f"poisson/A_{i}.ell" for i in range(1, 7)]
    + [f"poisson/prol_{i}.ell" for i in range(2, 7)]
    + ["poisson/rhs_6", "poisson/init_6", "poisson/sol_6"]
    # hierarchy used by math/mg_TEST.cc:223-387 (level 4)
    + [f"poisson_advanced4/sort_0/A_{i}.ell" for i in range(1, 5)]
    + [f"poisson_advanced4/sort_0/prol_{i}.ell" for i in range(2, 5)]
    + ["poisson_advanced4/sort_0/rhs_4", "poisson_advanced4/sort_0/init_4", "poisson_advanced4/sort_0/sol_4"]
    # PCG-Jacobi ELL system and FEATFLOW2 solution of math/dune_regression_TEST.cc:66-178
    + ["l2/area51_full_0.ell", "l2/area51_rhs_0", "l2/area51_init_0", "l2/area51_sol_0"]
)

for f in FILES:
    d = os.path.join(DST, f)
    os.makedirs(os.path.dirname(d), exist_ok=True)
    shutil.copyfile(os.path.join(SRC, f), d)
print(f"copied {len(FILES)} files to {DST}")
