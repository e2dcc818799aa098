"""Dense-matrix restatement of the Bloch Hamiltonian for small grids (test helper).

Built from textbook pieces, independently of the oracle's loops:
  H_k = -1/2 sum_d D2_d + |k+A|^2/2 + V - i sum_d (k+A)_d D1_d + sum_lma |p> rinv <p|
with D2_d / D1_d the circulant 8th-order finite-difference second / first
derivative matrices (coefficients cN(n,4)/h^2 and bN(n,4)/h), and the
projector |p_lma>(r_j) = zekr_uV(j, lma, k) at the sphere points.
Vectors are the interior (nz, ny, nx) grid flattened in C order.
"""
import numpy as np

from salmon2_amd.system import BN4, CN4


def circulant(n, coefs, antisym=False):
    M = np.zeros((n, n))
    for i in range(n):
        for k, c in enumerate(coefs):
            if k == 0:
                M[i, i] += c
                continue
            M[i, (i + k) % n] += c
            M[i, (i - k) % n] += -c if antisym else c
    return M


def kron3(Az, Ay, Ax):
    return np.kron(Az, np.kron(Ay, Ax))


def local_hamiltonian(s, kAc, ispin=0):
    n = (s.nx, s.ny, s.nz)
    I = [np.eye(m) for m in n]
    D2 = [circulant(n[d], [CN4[0] / s.hgs[d] ** 2] + [CN4[m] / s.hgs[d] ** 2 for m in range(1, 5)]) for d in range(3)]
    D1 = [circulant(n[d], [0.0] + [BN4[m] / s.hgs[d] for m in range(1, 5)], antisym=True) for d in range(3)]
    lap = kron3(I[2], I[1], D2[0]) + kron3(I[2], D2[1], I[0]) + kron3(D2[2], I[1], I[0])
    grad = [kron3(I[2], I[1], D1[0]), kron3(I[2], D1[1], I[0]), kron3(D1[2], I[1], I[0])]
    H = -0.5 * lap.astype(np.complex128)
    H += np.diag(s.vlocal[ispin].ravel()).astype(np.complex128)
    H += 0.5 * float(np.dot(kAc, kAc)) * np.eye(H.shape[0])
    for d in range(3):
        H += -1j * kAc[d] * grad[d]
    return H


def projector_matrix(s, zekr_k):
    """P (N_grid x nlma) with P[r_j, lma] = zekr_uV(j, lma, k)."""
    N = s.nx * s.ny * s.nz
    P = np.zeros((N, s.nlma), dtype=np.complex128)
    for l in range(s.nlma):
        ia = s.ia_tbl[l] - 1
        for j in range(s.mps[ia]):
            x, y, z = (int(v) - 1 for v in s.jxyz[ia, j])
            P[x + s.nx * (y + s.ny * z), l] += zekr_k[l, j]
    return P


def zekr_closed_form(s, kAc):
    """conj(exp(i (k+A).r)) uv for every (lma, j) (update_kvector_nonlocalpt)."""
    z = np.zeros((s.nlma, s.nps), dtype=np.complex128)
    for l in range(s.nlma):
        ia = s.ia_tbl[l] - 1
        m = s.mps[ia]
        th = s.rxyz[ia, :m] @ np.asarray(kAc)
        z[l, :m] = np.exp(-1j * th) * s.uv[l, :m]
    return z


def full_hamiltonian(s, kAc, ispin=0):
    H = local_hamiltonian(s, kAc, ispin)
    if s.natom:
        P = projector_matrix(s, zekr_closed_form(s, kAc))
        H += P @ np.diag(s.rinv_uvu) @ P.conj().T
    return H


def interior(psi_block, s):
    """(nz, ly, nx) padded orbital -> flat interior vector."""
    return np.ascontiguousarray(psi_block[:, : s.ny, :]).ravel()
