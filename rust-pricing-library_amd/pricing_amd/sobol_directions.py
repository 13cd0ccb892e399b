"""Joe--Kuo 6.21201 Sobol direction asset.

Writes the 21,201-dimension, 32-bit direction matrix in the reference's
``JK621201`` format (crates/pricing/data/README.md: 8-byte magic, big-endian
format version 1, dimension count 21201, bit count 32, then big-endian u32
direction words, dimension-major).  The source is SciPy's
``_sobol_direction_numbers.npz`` (primitive polynomials and initial direction
numbers of the same Joe--Kuo set the reference embeds); the directions follow
the Bratley--Fox / Joe--Kuo recurrence

    m_i = m_{i-s} ^ (m_{i-s} << s) ^ XOR_{k=1}^{s-1} a_k m_{i-k} << k,
    v_i = m_i << (31 - i),

with dimension 0 the identity (v_i = 2^(31-i)).  The resulting file is pinned
by its SHA-256 digest, which must equal the digest the reference records for
its embedded asset.
"""
import hashlib
import os
import struct

import numpy as np

DIMENSIONS = 21201
BITS = 32
SHA256 = "189f65c4e4fcf7455efb7618f3dafbbbaf70303fc35ecd380fa28a67ab896900"


def _scipy_npz_path():
    import scipy.stats

    return os.path.join(os.path.dirname(scipy.stats.__file__), "_sobol_direction_numbers.npz")


def direction_matrix():
    data = np.load(_scipy_npz_path())
    poly = data["poly"]
    vinit = data["vinit"]
    out = np.zeros((DIMENSIONS, BITS), dtype=np.uint64)
    out[0, :] = [1 << (31 - k) for k in range(BITS)]
    for j in range(1, DIMENSIONS):
        p = int(poly[j])
        s = p.bit_length() - 1
        m = [int(vinit[j, k]) for k in range(s)]
        for i in range(s, BITS):
            v = m[i - s] ^ (m[i - s] << s)
            for k in range(1, s):
                if (p >> (s - k)) & 1:
                    v ^= m[i - k] << k
            m.append(v)
        out[j, :] = [m[k] << (31 - k) for k in range(BITS)]
    if int(out.max()) >= 1 << 32:
        raise ValueError("direction word overflow")
    return out.astype(np.uint32)


def asset_bytes():
    words = direction_matrix()
    return b"JK621201" + struct.pack(">III", 1, DIMENSIONS, BITS) + words.astype(">u4").tobytes()


def write(path):
    data = asset_bytes()
    digest = hashlib.sha256(data).hexdigest()
    if digest != SHA256:
        raise RuntimeError(f"direction asset digest {digest} != {SHA256}")
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)
    return digest


def default_path():
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(os.path.dirname(here), "data", "joe_kuo_6_21201_u32be.bin")


if __name__ == "__main__":
    print(write(default_path()))
