"""Copies the marker positions of the reference's sphere-drag deck
(tests/drag_forces_sphere/sphere_markers.data: 162 vertices of a
frequency-4 icosahedral sphere of radius 3 about the origin, a LAMMPS data
file) into sphere_markers.npz.  Run in a container that holds the
reference: python tests/golden/make_sphere_markers.py"""
import os

import numpy as np

SRC = "/root/reference/tests/drag_forces_sphere/sphere_markers.data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere_markers.npz")


def main():
    text = open(SRC).read()
    atoms = text.split("Atoms # atomic", 1)[1].strip().splitlines()
    rows = [ln.split() for ln in atoms if ln.strip()]
    ids = np.array([int(r[0]) for r in rows])
    x = np.array([[float(r[2]), float(r[3]), float(r[4])] for r in rows])
    order = np.argsort(ids)
    np.savez(OUT, x=x[order], ids=ids[order])
    print(f"wrote {OUT}: {len(ids)} markers")


if __name__ == "__main__":
    main()
