"""Copy the reference's data fixtures for this path into tests/golden/.

Provenance (all data files of graham-g-brown/ggbTDDFT, read as data):
  src/workingData/mesh/legendreRoots/roots{00032,00064,00128,00256,00512}.txt
      interior Gauss-Lobatto nodes (roots of P'_{N+1})
  src/workingData/stateParameters/clebschGordan/lo00{0,2}.txt
      m = 0 Clebsch-Gordan coefficients <l1 0 l2 0|L 0> as rows "l1 l2 L value"
  src/workingData/stateParameters/densityFrozen.bin
      float64 x 512, the Xe core density in the reference's u-units / (4 pi)
Run once in a container where /root/reference exists.
"""
import os
import shutil

SRC = "/root/reference/src/workingData"
DST = os.path.dirname(os.path.abspath(__file__))
FILES = ["mesh/legendreRoots/roots%05d.txt" % n for n in (32, 64, 128, 256, 512)] + [
    "stateParameters/clebschGordan/lo000.txt",
    "stateParameters/clebschGordan/lo002.txt",
    "stateParameters/densityFrozen.bin",
]

if __name__ == "__main__":
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, os.path.basename(f)))
        print("copied", f)
