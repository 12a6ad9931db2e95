import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

MILC = dict(one_link=5.0 / 8, three_staple=-1.0 / 16, five_staple=1.0 / 64,
            seven_staple=-1.0 / 384, lepage=-1.0 / 16, naik=-1.0 / 24)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from _oracle import Oracle
    return Oracle()


def rand_su3(n, rng):
    """n Haar-random SU(3) matrices as (n,3,3) complex (QR + phase fix)."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q[:, :, 2] *= np.conj(det)[:, None] / np.abs(det)[:, None]
    return q


def to_raw(m):
    """(n,3,3) complex -> (n,3,3,2) float64."""
    return np.ascontiguousarray(np.stack([m.real, m.imag], axis=-1))


def from_raw(a):
    return a[..., 0] + 1j * a[..., 1]


def gauge_numpy(L, seed, eps=0.3):
    """Independent numpy stand-in for seeded near-identity SU(3) links."""
    rng = np.random.default_rng(seed)
    V = int(np.prod(L))
    U = []
    for mu in range(4):
        g = (rng.standard_normal((V, 3, 3)) + 1j * rng.standard_normal((V, 3, 3)))
        h = 0.5 * (g + np.conj(np.transpose(g, (0, 2, 1))))
        h -= np.trace(h, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
        w, v = np.linalg.eigh(h)
        u = np.einsum("nij,nj,nkj->nik", v, np.exp(1j * eps * w), np.conj(v))
        U.append(to_raw(u))
    return U
