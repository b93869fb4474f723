"""Writes tests/golden/julius_mesh.npz from the reference's own data file Data/Julius/julius.ply
(5,076 vertices, 9,999 triangles, per-face texture coordinates), read through the library's PLY
reader.  Run in a container that has the reference checked out; the GPU tests only read the npz."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
import gradientdomaintextureprocessing_amd as g  # noqa: E402

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Data/Julius/julius.ply"
v, t, uv = g.TexturedMesh.read_ply(src).arrays()
np.savez_compressed(os.path.join(HERE, "julius_mesh.npz"), vertices=v, triangles=t.astype(np.int32), corner_uv=uv)
print("wrote", v.shape, t.shape, uv.shape)
