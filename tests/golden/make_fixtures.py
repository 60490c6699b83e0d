"""Write the golden fixtures used by the CPU and GPU parity tests.

    python tests/golden/make_fixtures.py [/path/to/reference]

Reads DATA files of the reference checkout (never its code):
  benchmark_data/regular/heisenberg/L12_D20/{mpo.npz, initial_mps.npz, golden_results.json}
  benchmark_data/regular/josephson/L20_D50_nmax2/{mpo.npz, golden_results.json}
  benchmarks/data/{heisenberg_L12_mpo, heisenberg_L12_chi10_mps,
                   josephson_L8_n2_mpo, josephson_L8_n2_chi10_mps}.bin (copied as they are)
The npz files are opened with allow_pickle=False; their pickled 'metadata'
entry is skipped.  Tensors are converted from the quimb layout of those files
(MPO bulk [l, r, up, down], 3-D edges, left boundary at l = D-1, right boundary
at r = 0; MPS bulk (l, p, r), 2-D edges (p, r) and (l, p)) to this engine's
layout (MPO W[w, s, sp, wp], left boundary w = 0, right boundary wp = D-1;
MPS (cL, d, cR)), and written as plain arrays to
  tests/golden/heis_L12_D20.npz
  tests/golden/josephson_L20_D50_nmax2.npz
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def quimb_mpo_to_engine(z, L):
    """Bond index flipped (quimb starts at D-1 and ends at 0; the engine starts
    at 0 and ends at D-1); physical pair (up, down) -> (sp, s)."""
    bulk = z["tensor_1"]
    D, d = bulk.shape[0], bulk.shape[2]
    Ws = []
    for i in range(L):
        t = z["tensor_%d" % i]
        W = np.zeros((D, d, d, D), dtype=t.dtype)
        if i == 0:        # (r, up, down): the row l = D-1 of the bulk
            for r in range(D):
                W[0, :, :, D - 1 - r] = t[r].T
        elif i == L - 1:  # (l, up, down): the column r = 0
            for l in range(D):
                W[D - 1 - l, :, :, D - 1] = t[l].T
        else:
            for l in range(D):
                for r in range(D):
                    W[D - 1 - l, :, :, D - 1 - r] = t[l, r].T
        Ws.append(W)
    return np.stack(Ws)


def quimb_mps_to_engine(z, L):
    out = []
    for i in range(L):
        t = z["tensor_%d" % i]
        if i == 0:
            t = t[None, :, :]          # (p, r) -> (1, p, r)
        elif i == L - 1:
            t = t[:, :, None]          # (l, p) -> (l, p, 1)
        out.append(np.asarray(t))
    return out


def main(ref):
    base = os.path.join(ref, "benchmark_data", "regular")
    # Heisenberg L=12, chi=20
    d0 = os.path.join(base, "heisenberg", "L12_D20")
    z = np.load(os.path.join(d0, "mpo.npz"), allow_pickle=False)
    m = np.load(os.path.join(d0, "initial_mps.npz"), allow_pickle=False)
    g = json.load(open(os.path.join(d0, "golden_results.json")))
    L = 12
    W = quimb_mpo_to_engine(z, L)
    mps = quimb_mps_to_engine(m, L)
    arrs = {"W": W, "E_dmrg1": g["quimb_dmrg1"]["energy"], "E_dmrg2": g["quimb_dmrg2"]["energy"],
            "chi": 20}
    for i, A in enumerate(mps):
        arrs["mps_%d" % i] = A
    np.savez(os.path.join(HERE, "heis_L12_D20.npz"), **arrs)
    # Josephson L=20, chi=50, n_max=2, E_J=1, E_C=0.5, flux pi/4
    d1 = os.path.join(base, "josephson", "L20_D50_nmax2")
    z = np.load(os.path.join(d1, "mpo.npz"), allow_pickle=False)
    g = json.load(open(os.path.join(d1, "golden_results.json")))
    W = quimb_mpo_to_engine(z, 20)
    np.savez(os.path.join(HERE, "josephson_L20_D50_nmax2.npz"), W=W,
             E_dmrg1=g["quimb_dmrg1"]["energy"], E_dmrg2=g["quimb_dmrg2"]["energy"], chi=50)
    # .bin files of benchmarks/data (seed 42): copied byte for byte
    src = os.path.join(ref, "benchmarks", "data")
    for name in ("heisenberg_L12_mpo.bin", "heisenberg_L12_chi10_mps.bin",
                 "josephson_L8_n2_mpo.bin", "josephson_L8_n2_chi10_mps.bin"):
        with open(os.path.join(src, name), "rb") as fi, open(os.path.join(HERE, name), "wb") as fo:
            fo.write(fi.read())
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
