"""Copies the reference's smallest sample mesh (sample_meshes/
naca012_2D_triangles.meshb.xz: a 2D triangle mesh of 36,624 vertices, 465
boundary edges and 72,783 triangles, libMeshb version 1) into this folder
as a fixture for the mesh-file reader.  Run from the repository root with
the reference tree at /root/reference."""
import os
import shutil

SRC = "/root/reference/sample_meshes/naca012_2D_triangles.meshb.xz"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "naca012_2D_triangles.meshb.xz")

if __name__ == "__main__":
    shutil.copyfile(SRC, DST)
    print("wrote", DST, os.path.getsize(DST), "bytes")
