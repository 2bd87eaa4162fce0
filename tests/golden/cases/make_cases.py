#!/usr/bin/env python3
"""Copies the input files of the reference's own test cases that this path runs into fixtures beside this script
(input.ini and INPUT/*; the large files gzip-compressed, which mosegpu.caseio reads as they are).  Run in a
container where the reference tree is at /root/reference; the GPU tests read only the copies.

  2D/euler/isentropic-vortex   Euler, HLLC + van Leer, SSP-RK3, CFL 0.5, 64 x 64 x 1, periodic (MOSE 201) in x and y,
                               no-flux faces (MOSE 0) in the one-cell direction
  1D/Shu-Osher                 Euler, HLLC + van Leer, SSP-RK3, CFL 0.5, 200 x 1 x 1 on [-5, 5], extrapolation (MOSE 400)
                               in x, no-flux faces (MOSE 0) in y and z: a Mach 3 shock into a density sine wave
  1D/Ein91                     Euler, HLLE + van Leer, SSP-RK3, CFL 0.9, Einfeldt's 1-2-0-3 double rarefaction
  1D/Noh87                     Euler, HLLE, SSP-RK3, CFL 0.5, 400 cells: two cold streams colliding at the centre
  1D/Sod79                     Euler, SLAU + MC limiter, SSP-RK3, CFL 0.9, 400 x 1 x 1, extrapolation (MOSE 400) on
                               every face: Sod's shock tube to t = 0.2
  1D/diff1                     Navier-Stokes, 4 species (CH4, H2O, N2, O2; Schmidt 1, Prandtl 1), HLLC + van Leer,
                               SSP-RK2, CFL and VNN 0.95, 200 x 1 x 1, periodic (MOSE 201) in x, symmetry (MOSE 300)
                               in y, no-flux faces (MOSE 0) in z: species layers at rest that diffuse
"""
import gzip
import os
import shutil

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"isentropic-vortex": "2D/euler/isentropic-vortex", "diff1": "1D/diff1", "shu-osher": "1D/Shu-Osher", "ein91": "1D/Ein91", "noh87": "1D/Noh87",
         "sod79": "1D/Sod79", "forward-step": "2D/euler/supersonic-forward-step"}
COMPRESS = ("bc.txt", "ic.tec", "thermo.dat", "transport.dat")


def main():
    for name, rel in CASES.items():
        src, dst = os.path.join(REF, rel), os.path.join(HERE, name)
        os.makedirs(os.path.join(dst, "INPUT"), exist_ok=True)
        shutil.copyfile(os.path.join(src, "input.ini"), os.path.join(dst, "input.ini"))
        for f in sorted(os.listdir(os.path.join(src, "INPUT"))):
            a = os.path.join(src, "INPUT", f)
            if f in COMPRESS:
                with open(a, "rb") as fi, gzip.GzipFile(os.path.join(dst, "INPUT", f + ".gz"), "wb", 9, mtime=0) as fo:
                    shutil.copyfileobj(fi, fo)
            else:
                shutil.copyfile(a, os.path.join(dst, "INPUT", f))
        print(name, sorted(os.listdir(os.path.join(dst, "INPUT"))))


if __name__ == "__main__":
    main()
