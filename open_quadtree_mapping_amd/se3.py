"""Rigid transforms as 3x4 row-major float32 matrices [R | t].

The products and inverses follow the operation order of the reference's
SE3<float> (include/quadmap/se3.cuh:90-107 for inv(), :177-195 for the product),
evaluated in float32, so the relative poses handed to the kernels are the ones
the reference host code would compute from the same absolute poses.
"""
from __future__ import annotations

import numpy as np

__all__ = ["se3_from_rt", "se3_inv", "se3_mul", "rot_xyz", "relative_pose", "se3_from_quat"]

f32 = np.float32


def rot_xyz(ax: float, ay: float, az: float) -> np.ndarray:
    """Rotation Rz(az) @ Ry(ay) @ Rx(ax), float64."""
    cx, sx = np.cos(ax), np.sin(ax)
    cy, sy = np.cos(ay), np.sin(ay)
    cz, sz = np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def se3_from_rt(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    T = np.zeros((3, 4), dtype=np.float32)
    T[:, :3] = R
    T[:, 3] = t
    return T


def se3_from_quat(qw, qx, qy, qz, tx, ty, tz) -> np.ndarray:
    """SE3(qw, qx, qy, qz, tx, ty, tz) of include/quadmap/se3.cuh:48-77, in float32."""
    qw, qx, qy, qz = f32(qw), f32(qx), f32(qy), f32(qz)
    x, y, z = f32(2) * qx, f32(2) * qy, f32(2) * qz
    wx, wy, wz = x * qw, y * qw, z * qw
    xx, xy, xz = x * qx, y * qx, z * qx
    yy, yz, zz = y * qy, z * qy, z * qz
    T = np.zeros((3, 4), dtype=np.float32)
    T[0] = [f32(1) - (yy + zz), xy - wz, xz + wy, tx]
    T[1] = [xy + wz, f32(1) - (xx + zz), yz - wx, ty]
    T[2] = [xz - wy, yz + wx, f32(1) - (xx + yy), tz]
    return T


def se3_inv(T: np.ndarray) -> np.ndarray:
    d = np.asarray(T, dtype=np.float32).reshape(12)
    r = np.zeros(12, dtype=np.float32)
    r[0], r[1], r[2] = d[0], d[4], d[8]
    r[4], r[5], r[6] = d[1], d[5], d[9]
    r[8], r[9], r[10] = d[2], d[6], d[10]
    r[3] = ((-d[0]) * d[3] - d[4] * d[7]) - d[8] * d[11]
    r[7] = ((-d[1]) * d[3] - d[5] * d[7]) - d[9] * d[11]
    r[11] = ((-d[2]) * d[3] - d[6] * d[7]) - d[10] * d[11]
    return r.reshape(3, 4)


def se3_mul(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    a = np.asarray(A, dtype=np.float32).reshape(12)
    b = np.asarray(B, dtype=np.float32).reshape(12)
    r = np.zeros(12, dtype=np.float32)
    for row in range(3):
        o = 4 * row
        for col in range(3):
            r[o + col] = (a[o] * b[col] + a[o + 1] * b[4 + col]) + a[o + 2] * b[8 + col]
        r[o + 3] = ((a[o + 3] + a[o] * b[3]) + a[o + 1] * b[7]) + a[o + 2] * b[11]
    return r.reshape(3, 4)


def relative_pose(T_ref_world: np.ndarray, T_cur_world: np.ndarray) -> np.ndarray:
    """T_ref_cur = T_ref_world * T_cur_world^-1 (src/seed_matrix.cu:462)."""
    return se3_mul(T_ref_world, se3_inv(T_cur_world))
