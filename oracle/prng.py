"""Counter-based splitmix64 generator (oracle, test-only).

The reference seeds its random MPS from the global C ``rand()``
(gpu-rocm/common/scalar_traits.h:62; dmrg2_gpu_impl.h:382-393) and reseeds it
with ``srand(42 + site)`` for a zero-norm Lanczos start (:983-989).  That is not
reproducible across implementations, so the engine defines its own generator:
output ``k`` of stream ``seed`` is ``splitmix64(seed + (k+1)*0x9E3779B97F4A7C15)``,
mapped to ``u = (z >> 11) * 2**-53`` in [0, 1) and then to ``2u - 1``.
This module restates it bit-exactly in numpy so the oracle starts from the
same MPS as the engine.

Pinning: the first outputs of splitmix64 for seed 0 are the published test
values of the algorithm (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, ...), checked in
tests/test_oracle_models.py; the C implementation is compared bit-exactly on
the GPU through ``dmrg_init_mps_random``.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """Return outputs offset..offset+count-1 of the stream as uint64."""
    with np.errstate(over='ignore'):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed, count, offset=0):
    """Doubles in [-1, 1): 2 * ((z >> 11) * 2^-53) - 1."""
    z = splitmix64(seed, count, offset)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def random_tensor_stream(seed, sizes, complex_=False, scale=1.0):
    """Split one stream into consecutive chunks (Fortran-order tensors).

    Complex entries take two consecutive outputs (real, imaginary)."""
    per = 2 if complex_ else 1
    total = per * int(sum(sizes))
    u = uniform_pm1(seed, total) * scale
    out, pos = [], 0
    for n in sizes:
        chunk = u[pos:pos + per * n]
        pos += per * n
        if complex_:
            chunk = chunk[0::2] + 1j * chunk[1::2]
        out.append(chunk)
    return out
