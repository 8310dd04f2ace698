"""Copies the reference tests' own 3-D gmsh fixtures (data files held under
tests/morganaMeshes of the morganaPublic tree) into tests/golden/meshes, so the
parity tests can assemble the meshes the reference's assembly tests load
(e.g. mignon3dB.msh in tests/morganaFiniteElements/traitsMatrixBuilderTest5.cpp).
Run here, where the reference tree is mounted."""
import glob
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes")
os.makedirs(OUT, exist_ok=True)
names = sorted(glob.glob(os.path.join(REF, "tests", "morganaMeshes", "mignon3d*.msh")))
names.append(os.path.join(REF, "tests", "morganaMeshes", "intBoundary3dA.msh"))
for src in names:
    shutil.copyfile(src, os.path.join(OUT, os.path.basename(src)))
    print("copied", os.path.basename(src))
