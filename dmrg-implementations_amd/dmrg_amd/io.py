"""Reader and writer for the reference's benchmark data files.

Two formats, both data only (nothing is executed or unpickled):

* ``.bin`` (benchmarks/lib/data_loader.py:24-88, written by
  benchmarks/data/generate.py:24-103): int64 ``L``, int64 bond dimensions
  ``[L+1]``, int64 physical dimensions ``[L]``, then per site the int64 shape
  (3 entries for an MPS tensor, 4 for an MPO tensor) followed by the complex128
  entries in C order.  MPS tensors are (Dl, d, Dr).  MPO tensors are
  (Dl, up, down, Dr) in quimb's bond convention: the bulk starts at bond index
  D-1 on the left and ends at 0 on the right, and the edge tensors have a
  bond of size 1 on the outside.

* the regular-tier npz fixtures (``benchmark_data/regular/*``): quimb tensors
  ``tensor_0 .. tensor_{L-1}``; MPO bulk [l, r, up, down] with 3-D edges, MPS
  bulk (l, p, r) with 2-D edges.  Opened with ``allow_pickle=False``, so the
  pickled ``metadata`` entry is never loaded.

Engine layouts (gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:22-34): MPO
``W[w, s, sp, wp]`` (numpy (D, d, d, D); the C-ABI takes it in Fortran order)
with s the ket index, left boundary row w = 0 and right boundary column
wp = D-1; MPS (cL, d, cR).  The importers flip the bond index (w -> D-1-w)
and map the physical pair (up, down) -> (sp, s).  For the Hermitian models of
the reference a transposed physical pair maps H to H^T = conj(H), which has
the same spectrum; the ket/bra order itself is pinned by <psi|H|psi> on a
complex MPS in tests/test_io.py.
"""
import numpy as np

__all__ = ["read_bin", "write_bin", "mpo_bin_to_engine", "mpo_engine_to_bin", "mps_bin_to_engine",
           "mps_engine_to_bin", "load_bin_case", "load_npz_mpo", "load_npz_mps"]


def read_bin(path, kind):
    """Tensors of a ``.bin`` file; kind "mps" (3-D) or "mpo" (4-D)."""
    if kind not in ("mps", "mpo"):
        raise ValueError("kind must be 'mps' or 'mpo'")
    nd = 3 if kind == "mps" else 4
    with open(path, "rb") as f:
        L = int(np.fromfile(f, np.int64, 1)[0])
        if L < 1 or L > 1 << 20:
            raise ValueError("%s: bad site count %d" % (path, L))
        bonds = np.fromfile(f, np.int64, L + 1)
        phys = np.fromfile(f, np.int64, L)
        if bonds.size != L + 1 or phys.size != L:
            raise ValueError("%s: truncated header" % path)
        out = []
        for i in range(L):
            shape = tuple(int(x) for x in np.fromfile(f, np.int64, nd))
            if len(shape) != nd:
                raise ValueError("%s: truncated at site %d" % (path, i))
            want = (int(bonds[i]),) + (int(phys[i]),) * (nd - 2) + (int(bonds[i + 1]),)
            if shape != want:
                raise ValueError("%s: site %d shape %s, header says %s" % (path, i, shape, want))
            n = int(np.prod(shape))
            data = np.fromfile(f, np.complex128, n)
            if data.size != n:
                raise ValueError("%s: truncated data at site %d" % (path, i))
            out.append(data.reshape(shape))
        if f.read(1):
            raise ValueError("%s: trailing bytes" % path)
    return out


def write_bin(path, tensors):
    """Write tensors (all 3-D or all 4-D, bonds chained) in the ``.bin`` layout."""
    ts = [np.asarray(t, dtype=np.complex128) for t in tensors]
    nd = ts[0].ndim
    if nd not in (3, 4) or any(t.ndim != nd for t in ts):
        raise ValueError("tensors must all be 3-D (MPS) or 4-D (MPO)")
    for i in range(len(ts) - 1):
        if ts[i].shape[-1] != ts[i + 1].shape[0]:
            raise ValueError("bond mismatch between sites %d and %d" % (i, i + 1))
    with open(path, "wb") as f:
        np.array([len(ts)], np.int64).tofile(f)
        np.array([ts[0].shape[0]] + [t.shape[-1] for t in ts], np.int64).tofile(f)
        np.array([t.shape[1] for t in ts], np.int64).tofile(f)
        for t in ts:
            np.array(t.shape, np.int64).tofile(f)
            np.ascontiguousarray(t).tofile(f)


def _real_if_possible(arrs):
    if all(not np.any(np.imag(a)) for a in arrs):
        return [np.ascontiguousarray(np.real(a)) for a in arrs], "f64"
    return [np.asarray(a, dtype=np.complex128) for a in arrs], "c128"


def mpo_bin_to_engine(tensors):
    """(Dl, up, down, Dr) quimb-convention tensors -> engine W[w, s, sp, wp].

    Returns (list of (D, d, d, D) arrays, dtype "f64" or "c128")."""
    L = len(tensors)
    D = max(max(t.shape[0], t.shape[3]) for t in tensors)
    Ws = []
    for i, t in enumerate(tensors):
        dl, d, d2, dr = t.shape
        if d != d2:
            raise ValueError("site %d: physical dimensions %d, %d differ" % (i, d, d2))
        W = np.zeros((D, d, d, D), dtype=np.complex128)
        # an outer edge bond of size 1 is the boundary row/column of the bulk
        lmap = (lambda l: 0) if (i == 0 and dl == 1) else (lambda l: D - 1 - l)
        rmap = (lambda r: D - 1) if (i == L - 1 and dr == 1) else (lambda r: D - 1 - r)
        for l in range(dl):
            for r in range(dr):
                W[lmap(l), :, :, rmap(r)] = t[l, :, :, r].T
        Ws.append(W)
    return _real_if_possible(Ws)


def mpo_engine_to_bin(Ws):
    """Inverse of mpo_bin_to_engine: engine W[w, s, sp, wp] -> (Dl, up, down, Dr)
    with 1-wide outer edge bonds."""
    L = len(Ws)
    out = []
    for i, W in enumerate(Ws):
        D, d = W.shape[0], W.shape[1]
        ls = [0] if i == 0 else list(range(D))
        rs = [D - 1] if i == L - 1 else list(range(D))
        t = np.zeros((len(ls), d, d, len(rs)), dtype=np.complex128)
        for a, w in enumerate(ls):
            for b, wp in enumerate(rs):
                la = 0 if i == 0 else D - 1 - w
                rb = 0 if i == L - 1 else D - 1 - wp
                t[la, :, :, rb] = W[w, :, :, wp].T
        out.append(t)
    return out


def mps_bin_to_engine(tensors):
    """(Dl, d, Dr) tensors are already the engine's (cL, d, cR); returns
    (list, dtype "f64" or "c128")."""
    return _real_if_possible([np.asarray(t) for t in tensors])


def mps_engine_to_bin(As):
    return [np.asarray(a, dtype=np.complex128) for a in As]


def load_bin_case(mps_path, mpo_path):
    """(Ws, As, dtype) ready for DMRG.set_mpo / DMRG.set_mps; dtype is "c128"
    when either file has a non-zero imaginary part, else "f64"."""
    Ws, tw = mpo_bin_to_engine(read_bin(mpo_path, "mpo"))
    As, ta = mps_bin_to_engine(read_bin(mps_path, "mps"))
    if len(Ws) != len(As):
        raise ValueError("MPO has %d sites, MPS %d" % (len(Ws), len(As)))
    if tw == "c128" or ta == "c128":
        return ([np.asarray(W, np.complex128) for W in Ws], [np.asarray(A, np.complex128) for A in As],
                "c128")
    return Ws, As, "f64"


def _npz_tensors(path):
    with np.load(path, allow_pickle=False) as z:
        keys = sorted((k for k in z.files if k.startswith("tensor_")), key=lambda k: int(k[7:]))
        return [np.asarray(z[k]) for k in keys]


def load_npz_mpo(path):
    """Regular-tier npz MPO (bulk [l, r, up, down], 3-D edges (r, up, down) and
    (l, up, down)) -> (engine Ws, dtype)."""
    ts = _npz_tensors(path)
    L = len(ts)
    conv = []
    for i, t in enumerate(ts):
        if i == 0 and t.ndim == 3:
            t = t[None]                      # (1, r, up, down)
        elif i == L - 1 and t.ndim == 3:
            t = t[:, None]                   # (l, 1, up, down)
        conv.append(np.transpose(t, (0, 2, 3, 1)))  # (l, up, down, r)
    return mpo_bin_to_engine(conv)


def load_npz_mps(path):
    """Regular-tier npz MPS (bulk (l, p, r), 2-D edges (p, r) and (l, p)) ->
    (engine As, dtype)."""
    ts = _npz_tensors(path)
    L = len(ts)
    out = []
    for i, t in enumerate(ts):
        if i == 0 and t.ndim == 2:
            t = t[None]
        elif i == L - 1 and t.ndim == 2:
            t = t[:, :, None]
        out.append(t)
    return mps_bin_to_engine(out)

