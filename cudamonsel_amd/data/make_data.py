"""Convert the NIST SRD-64 elastic tables and the atomic-weight list shipped with
CudaMONSEL into the compact files this package loads at run time.

Inputs (read-only, from the reference checkout):
  gov/nist/microanalysis/EPQLibrary/NistXSec/Ezz.D64   one number per line:
      Z, then 61 blocks of (sigma_j in Bohr-radius^2, 201 quantiles q_jk)
      (format parsed by NISTMottScatteringAngle::loadData,
       gov/nist/microanalysis/EPQLibrary/NISTMottScatteringAngle.cu:44-66)
  gov/nist/microanalysis/EPQLibrary/AtomicWeights.csv  one weight per line,
      '//' comment header (read by Element.cu:593-620 as float)

  gov/nist/microanalysis/EPQLibrary/FFastEdgeDB.csv  edge energies (eV) per
      element Z = 1..92, one row each, shells K, L1-3, M1-5, N1-7, O1-5,
      P1-3 (read by ChantlerEdgeEnergy::loadFFastEdgeDB, EdgeEnergy.cu:153-177,
      which keeps each row's non-zero values)
  gov/nist/microanalysis/EPQLibrary/ElectronConfig.csv  ground-state shell
      occupancies (ENDLIB97) per element, "Z, n_K, n_L1, ..." with '#'
      comments (AtomicShell.cu:386-412)

Outputs (written next to this script):
  nistmott.bin        little-endian: b"NMOT", u32 version=1, u32 count,
                      then per element: i32 Z, f64 spwem[61], f64 x1[61*201]
  atomic_weights.txt  one weight per line for Z = 1.., as text
  edge_energies.txt   line Z-1: the element's non-zero edge energies (eV)
  electron_config.txt line Z-1: the element's shell occupancies

Usage: python make_data.py [reference_root] [Z ...]
"""
import os
import struct
import sys

SPWEM_LEN = 61
X1_LEN = 201
DEFAULT_Z = [1, 6, 7, 8, 14, 29]
EPQ = "CudaMONSEL/CudaMONSEL/gov/nist/microanalysis/EPQLibrary"


def read_d64(path):
    with open(path) as f:
        vals = [float(line) for line in f if line.strip()]
    z = int(vals[0])
    expect = 1 + SPWEM_LEN * (1 + X1_LEN)
    if len(vals) != expect:
        raise ValueError(f"{path}: {len(vals)} numbers, expected {expect}")
    spwem, x1 = [], []
    pos = 1
    for _ in range(SPWEM_LEN):
        spwem.append(vals[pos])
        x1.extend(vals[pos + 1: pos + 1 + X1_LEN])
        pos += 1 + X1_LEN
    return z, spwem, x1


def main(argv):
    root = argv[1] if len(argv) > 1 else "/root/reference"
    zs = [int(a) for a in argv[2:]] or DEFAULT_Z
    here = os.path.dirname(os.path.abspath(__file__))
    out = bytearray(b"NMOT" + struct.pack("<II", 1, len(zs)))
    for z in zs:
        zz, spwem, x1 = read_d64(os.path.join(root, EPQ, "NistXSec", f"E{z:02d}.D64"))
        assert zz == z
        out += struct.pack("<i", z)
        out += struct.pack(f"<{SPWEM_LEN}d", *spwem)
        out += struct.pack(f"<{SPWEM_LEN * X1_LEN}d", *x1)
    with open(os.path.join(here, "nistmott.bin"), "wb") as f:
        f.write(out)
    weights = []
    with open(os.path.join(root, EPQ, "AtomicWeights.csv")) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("//"):
                continue
            weights.append(s.split(",")[0])
    with open(os.path.join(here, "atomic_weights.txt"), "w") as f:
        f.write("\n".join(weights) + "\n")
    edges = []
    with open(os.path.join(root, EPQ, "FFastEdgeDB.csv")) as f:
        for line in f:
            vals = [float(v) for v in line.strip().split(",") if v.strip()]
            edges.append(" ".join(repr(v) for v in vals if v != 0.0))
    with open(os.path.join(here, "edge_energies.txt"), "w") as f:
        f.write("\n".join(edges) + "\n")
    occ = []
    with open(os.path.join(root, EPQ, "ElectronConfig.csv")) as f:
        for line in f:
            if not line.strip() or line.lstrip().startswith("#"):
                continue
            vals = [int(v) for v in line.strip().split(",") if v.strip()]
            occ.append(" ".join(str(v) for v in vals[1:]))
    with open(os.path.join(here, "electron_config.txt"), "w") as f:
        f.write("\n".join(occ) + "\n")
    print(f"wrote {len(zs)} Mott tables, {len(weights)} atomic weights, {len(edges)} edge rows, "
          f"{len(occ)} occupancy rows")


if __name__ == "__main__":
    main(sys.argv)
