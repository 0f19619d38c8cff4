"""Rigid transforms in the reference's representation (host-side input forming).

Transforms are 16 doubles in OpenGL column-major order, what
btTransform::getOpenGLMatrix writes (3rdparty/LinearMath/btTransform.h:136-143).
"""
from __future__ import annotations

import numpy as np


def quat_to_gl(q: np.ndarray, origin: np.ndarray) -> np.ndarray:
    """btTransform(btQuaternion q, origin).getOpenGLMatrix for a batch.

    q: (..., 4) as (x, y, z, w); origin: (..., 3).  The basis follows
    btMatrix3x3::setRotation's scalar path (btMatrix3x3.h:273-280).
    """
    q = np.asarray(q, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    d = x * x + y * y + z * z + w * w
    s = 2.0 / d
    xs, ys, zs = x * s, y * s, z * s
    wx, wy, wz = w * xs, w * ys, w * zs
    xx, xy, xz = x * xs, x * ys, x * zs
    yy, yz, zz = y * ys, y * zs, z * zs
    m = np.zeros(q.shape[:-1] + (16,), dtype=np.float64)
    # column 0
    m[..., 0] = 1.0 - (yy + zz)
    m[..., 1] = xy + wz
    m[..., 2] = xz - wy
    # column 1
    m[..., 4] = xy - wz
    m[..., 5] = 1.0 - (xx + zz)
    m[..., 6] = yz + wx
    # column 2
    m[..., 8] = xz + wy
    m[..., 9] = yz - wx
    m[..., 10] = 1.0 - (xx + yy)
    m[..., 12:15] = origin
    m[..., 15] = 1.0
    return m


def identity_gl(n: int | None = None) -> np.ndarray:
    m = np.zeros(16)
    m[[0, 5, 10, 15]] = 1.0
    return m if n is None else np.tile(m, (n, 1))


def translation_gl(t) -> np.ndarray:
    m = identity_gl()
    m[12:15] = t
    return m


def basis(m: np.ndarray) -> np.ndarray:
    """Row-major 3x3 basis of a GL matrix (or a batch)."""
    m = np.asarray(m)
    return np.stack([m[..., 0:3], m[..., 4:7], m[..., 8:11]], axis=-1)


def gl_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """btTransform a * b (btTransform.h:231-235) for GL matrices."""
    A, B = basis(a), basis(b)
    R = A @ B
    t = np.einsum("...ij,...j->...i", A, np.asarray(b)[..., 12:15]) + np.asarray(a)[..., 12:15]
    out = np.zeros(np.broadcast_shapes(np.shape(a), np.shape(b)))
    out[..., 0:3] = R[..., :, 0]
    out[..., 4:7] = R[..., :, 1]
    out[..., 8:11] = R[..., :, 2]
    out[..., 12:15] = t
    out[..., 15] = 1.0
    return out


def random_unit_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's subgroup method), (x,y,z,w)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    t1, t2 = 2.0 * np.pi * u2, 2.0 * np.pi * u3
    return np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], axis=-1)
