"""Writes tests/golden/fixtures.npz: small known-answer vectors for the oracle.

Every expected value is computed here with scipy.linalg (expm, expm_frechet)
and explicit numpy products, independently of the C oracle:

  herm_*     exp(-i H dt) for random Hermitian H           (scipy expm)
  gen_*      exp(A) for non-normal complex A, incl. ||A|| > 5.4 (scaling)
  fre_*      Frechet derivative L(X, E)                      (scipy expm_frechet)
  gate_*     d=4, n_t=5, 2-member gate ensemble: F and the first-order GRAPE
             gradient of src/Physics/Gradients.jl:254-291
  pair_*     d=2 state transfer (MR state kernel, GRAPEState.jl:186-231)
  lind_*     one damped qubit (Liouville dim 4), 2 state pairs, :square
             (GRAPELindblad.jl:146-196)

Run: python tests/golden/gen_golden.py   (numpy + scipy only)
"""
import os

import numpy as np
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")


def rand_herm(d, rng, scale=1.0):
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return scale * (a + a.conj().T) / 2


def haar(d, rng):
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))[None, :]


def gate_reference(H0s, Hc, scales, u, dt, Ut):
    """Ensemble mean of F = |Tr(Ut^dag U)/d|^2 and of the first-order gradient
    G[j,k] = 2 Re(conj(Phi) (-i dt) Tr(P_k Ut^dag R_k H_j)/d), P_k = U_{k-1}..U_1,
    R_k = Q_k^dag = U_n..U_k: the backward propagators of
    src/Computation/Propagators.jl:375-380 include slice k's own U_k."""
    M = len(H0s)
    nc, nt = u.shape
    d = Ut.shape[0]
    Fs, Gs = [], []
    for m in range(M):
        Hs = [scales[m, j] * Hc[j] for j in range(nc)]
        Us = [sl.expm(-1j * dt * (H0s[m] + sum(u[j, k] * Hs[j] for j in range(nc)))) for k in range(nt)]
        P = [np.eye(d)]
        for k in range(nt):
            P.append(Us[k] @ P[-1])
        Phi = np.trace(Ut.conj().T @ P[-1]) / d
        G = np.zeros((nc, nt))
        for k in range(nt):
            R = np.eye(d)
            for kk in range(k, nt):
                R = Us[kk] @ R
            A = P[k] @ Ut.conj().T @ R
            for j in range(nc):
                G[j, k] = 2 * np.real(np.conj(Phi) * (-1j * dt * np.trace(A @ Hs[j]) / d))
        Fs.append(abs(Phi) ** 2)
        Gs.append(G)
    return np.mean(Fs), np.mean(Gs, axis=0)


def pairs_reference(Us_fn, Ops, u, dts, psi0, psit, square, closed):
    """MR-kernel F (sum over pairs) and gradient for one member."""
    nc, nt = u.shape
    Us = [Us_fn(k) for k in range(nt)]
    F = 0.0
    G = np.zeros((nc, nt))
    for s in range(psi0.shape[0]):
        psi = [psi0[s]]
        for k in range(nt):
            psi.append(Us[k] @ psi[-1])
        z = np.vdot(psit[s], psi[-1])
        F += abs(z) ** 2 if square else z.real
        lam = [None] * (nt + 1)
        lam[nt] = psit[s]
        for k in range(nt - 1, -1, -1):
            lam[k] = Us[k].conj().T @ lam[k + 1]
        for k in range(nt):
            for j in range(nc):
                inner = np.vdot(lam[k + 1], Ops[j] @ psi[k])
                if closed:
                    g = 2 * dts[k] * np.imag(np.conj(z) * inner) if square else dts[k] * np.imag(inner)
                else:
                    g = 2 * dts[k] * np.real(np.conj(z) * inner) if square else dts[k] * np.real(inner)
                G[j, k] += g
    return F, G


def main():
    rng = np.random.default_rng(20261203)
    out = {}
    for i, d in enumerate((2, 4, 8, 16)):
        H = rand_herm(d, rng, scale=10.0 ** rng.uniform(-1, 2))
        dt = float(10.0 ** rng.uniform(-3, -1))
        out[f"herm_{i}_H"] = H
        out[f"herm_{i}_dt"] = np.array(dt)
        out[f"herm_{i}_U"] = sl.expm(-1j * dt * H)
    for i, (n, sc) in enumerate(((4, 0.01), (8, 0.3), (8, 2.0), (16, 7.0), (64, 0.5), (64, 12.0))):
        A = sc * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
        A += np.triu(rng.standard_normal((n, n)), 1) * sc  # non-normal part
        out[f"gen_{i}_A"] = A
        out[f"gen_{i}_E"] = sl.expm(A)
    for i, n in enumerate((3, 6)):
        X = 0.7 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        out[f"fre_{i}_X"] = X
        out[f"fre_{i}_E"] = E
        out[f"fre_{i}_L"] = sl.expm_frechet(X, E, compute_expm=False)
    # gate ensemble
    d, nc, nt, M = 4, 3, 5, 2
    H0s = np.stack([rand_herm(d, rng, 3.0) for _ in range(M)])
    Hc = np.stack([rand_herm(d, rng, 1.0) for _ in range(nc)])
    scales = rng.uniform(0.8, 1.2, (M, nc))
    u = rng.uniform(-2, 2, (nc, nt))
    dt = 0.07
    Ut = haar(d, rng)
    F, G = gate_reference(H0s, Hc, scales, u, dt, Ut)
    out.update(gate_H0=H0s, gate_Hc=Hc, gate_scale=scales, gate_u=u, gate_dt=np.array(dt), gate_Ut=Ut,
               gate_F=np.array(F), gate_G=G)
    # closed state pair (MR kernel, :square, with per-slice dt)
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]], complex)
    sz = np.diag([1.0 + 0j, -1.0])
    H0 = 0.8 * sz
    Ops = np.stack([0.5 * sx, 0.5 * sy])
    pwr = 1.3
    nt = 4
    u = rng.uniform(-1, 1, (2, nt))
    dts = rng.uniform(0.1, 0.3, nt)
    psi0 = np.array([[1, 0], [0.6, 0.8j]], complex)
    psit = np.array([[0, 1], [0.8, -0.6]], complex)
    fn = lambda k: sl.expm(-1j * dts[k] * (H0 + pwr * sum(u[j, k] * Ops[j] for j in range(2))))
    F, G = pairs_reference(fn, pwr * Ops, u, dts, psi0, psit, True, True)
    out.update(pair_H0=H0, pair_Hc=Ops, pair_pwr=np.array(pwr), pair_u=u, pair_dts=dts, pair_psi0=psi0,
               pair_psit=psit, pair_F=np.array(F), pair_G=G)
    # Lindblad: one qubit, amplitude damping + dephasing, controls x/y
    I2 = np.eye(2)
    Lm = np.array([[0, 0], [1, 0]], complex)

    def liou(H):
        return -1j * np.kron(I2, H) + 1j * np.kron(H.T, I2)

    def diss(L, g):
        A = L.conj().T @ L
        return g * (np.kron(L.conj(), L) - 0.5 * np.kron(I2, A) - 0.5 * np.kron(A.conj(), I2))

    L0 = liou(0.9 * sz / 2) + diss(Lm, 0.3) + diss(sz, 0.1)
    Lc = np.stack([liou(0.5 * sx), liou(0.5 * sy)])
    nt = 4
    u = rng.uniform(-2, 2, (2, nt))
    dts = np.full(nt, 0.2)
    r0 = np.array([[1, 0], [0, 0]], complex)
    r1 = np.array([[0.5, 0.5], [0.5, 0.5]], complex)
    t0 = np.array([[0, 0], [0, 1]], complex)
    t1 = np.array([[0.5, -0.5j], [0.5j, 0.5]], complex)
    s0 = np.stack([r0.reshape(-1, order="F"), r1.reshape(-1, order="F")])
    st = np.stack([t0.reshape(-1, order="F"), t1.reshape(-1, order="F")])
    pw = 0.9
    fn = lambda k: sl.expm(dts[k] * (L0 + pw * sum(u[j, k] * Lc[j] for j in range(2))))
    F, G = pairs_reference(fn, pw * Lc, u, dts, s0, st, True, False)
    out.update(lind_L0=L0, lind_Lc=Lc, lind_pwr=np.array(pw), lind_u=u, lind_dts=dts, lind_s0=s0, lind_st=st,
               lind_F=np.array(F), lind_G=G)
    np.savez(OUT, **out)
    print("wrote", OUT, len(out), "arrays")


if __name__ == "__main__":
    main()
