"""Generate the golden vectors that pin the CPU oracle.

Independent of the reference and of the oracle: 40-digit mpmath direct sums
and sympy's exact differentiation, from the closed forms
  G_lap3 = 1/|d|,  G_helm3 = exp(i k |d|)/|d|,  G_lap2 = log|d|
(sumpy/kernel.py:648-658, :789-791; before the global scaling constants),
and the QBX line-Taylor sum sum_{i<=p} (1/i!) d^i/dtau^i G(a + tau b)|_0
(sumpy/expansion/local.py:132-185) by mpmath Taylor differentiation.

Run:  python tests/golden/make_golden.py   (writes tests/golden/*.npz)
"""

import os

import mpmath as mp
import numpy as np
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40


def p2p_fixture():
    rng = np.random.default_rng(2024)
    out = {}
    ns, nt = 24, 17
    for dim in (2, 3):
        src = rng.random((dim, ns))
        tgt = rng.random((dim, nt)) + 0.3
        q = rng.standard_normal(ns)
        wd = rng.standard_normal((dim, ns))
        pot = np.zeros(nt)
        potd = np.zeros(nt)
        grad = np.zeros((dim, nt))
        for t in range(nt):
            acc, accd = mp.mpf(0), mp.mpf(0)
            g = [mp.mpf(0)]*dim
            for s in range(ns):
                d = [mp.mpf(tgt[a, t]) - mp.mpf(src[a, s]) for a in range(dim)]
                r2 = sum(x*x for x in d)
                if dim == 3:
                    G = 1/mp.sqrt(r2)
                    gr = [-x/r2**mp.mpf(1.5) for x in d]
                else:
                    G = mp.log(r2)/2
                    gr = [x/r2 for x in d]
                acc += q[s]*G
                accd += sum(mp.mpf(wd[a, s])*gr[a] for a in range(dim))
                for a in range(dim):
                    g[a] += q[s]*gr[a]
            pot[t], potd[t] = float(acc), float(accd)
            grad[:, t] = [float(x) for x in g]
        out[f"lap{dim}_src"], out[f"lap{dim}_tgt"] = src, tgt
        out[f"lap{dim}_q"], out[f"lap{dim}_wd"] = q, wd
        out[f"lap{dim}_pot"], out[f"lap{dim}_potdip"], out[f"lap{dim}_grad"] = pot, potd, grad
    # Helmholtz 3D, k = 5
    src = rng.random((3, ns))
    tgt = rng.random((3, nt)) + 0.3
    q = rng.standard_normal(ns)
    k = mp.mpf(5)
    pot = np.zeros(nt, complex)
    grad = np.zeros((3, nt), complex)
    for t in range(nt):
        acc = mp.mpc(0)
        g = [mp.mpc(0)]*3
        for s in range(ns):
            d = [mp.mpf(tgt[a, t]) - mp.mpf(src[a, s]) for a in range(3)]
            r = mp.sqrt(sum(x*x for x in d))
            G = mp.exp(1j*k*r)/r
            dG = mp.exp(1j*k*r)*(1j*k*r - 1)/r**3
            acc += q[s]*G
            for a in range(3):
                g[a] += q[s]*dG*d[a]
        pot[t] = complex(acc)
        grad[:, t] = [complex(x) for x in g]
    out.update(helm3_src=src, helm3_tgt=tgt, helm3_q=q, helm3_k=5.0, helm3_pot=pot,
               helm3_grad=grad)
    np.savez(os.path.join(HERE, "p2p_direct.npz"), **out)


def derivative_fixture():
    """rho^|m| d^m (1/r) at rational points, all |m| <= 12 with m_x <= 2."""
    x, y, z = sp.symbols("x y z")
    f = 1/sp.sqrt(x**2 + y**2 + z**2)
    pts = [(sp.Rational(7, 10), sp.Rational(-2, 5), sp.Rational(11, 10)),
           (sp.Rational(-3, 2), sp.Rational(1, 4), sp.Rational(1, 3))]
    rscale = sp.Rational(3, 10)
    mis, vals = [], []
    for pi, p in enumerate(pts):
        for n in range(13):
            for mx in range(min(n, 2) + 1):
                for my in range(n - mx + 1):
                    mz = n - mx - my
                    e = f
                    if mx:
                        e = sp.diff(e, x, mx)
                    if my:
                        e = sp.diff(e, y, my)
                    if mz:
                        e = sp.diff(e, z, mz)
                    v = sp.N(e.subs({x: p[0], y: p[1], z: p[2]})*rscale**n, 30)
                    mis.append((pi, mx, my, mz))
                    vals.append(float(v))
    np.savez(os.path.join(HERE, "laplace3d_derivatives.npz"),
             points=np.array([[float(c) for c in p] for p in pts]), rscale=float(rscale),
             mis=np.array(mis, np.int32), values=np.array(vals))


def qbx_fixture():
    """Line-Taylor sums of 2D Laplace SLP / DLP and their a-derivatives."""
    rng = np.random.default_rng(7)
    npair = 12
    order = 6
    a = rng.standard_normal((2, npair))
    b = 0.3*rng.standard_normal((2, npair))
    n = rng.standard_normal((2, npair))
    n /= np.linalg.norm(n, axis=0)
    res = np.zeros((4, npair))  # SLP, DLP, d/da0 SLP, d/da1 DLP

    def line_sum(F, i):
        tay = mp.taylor(lambda tau: F(tau), 0, order)
        return sum(tay)

    for i in range(npair):
        A = [mp.mpf(a[0, i]), mp.mpf(a[1, i])]
        B = [mp.mpf(b[0, i]), mp.mpf(b[1, i])]
        N = [mp.mpf(n[0, i]), mp.mpf(n[1, i])]

        def G(x0, x1):
            return mp.log(x0*x0 + x1*x1)/2

        def slp(tau):
            return G(A[0] + tau*B[0], A[1] + tau*B[1])

        def dlp(tau, A=A):
            x0, x1 = A[0] + tau*B[0], A[1] + tau*B[1]
            r2 = x0*x0 + x1*x1
            return -(N[0]*x0 + N[1]*x1)/r2

        def dslp_da0(tau):
            x0, x1 = A[0] + tau*B[0], A[1] + tau*B[1]
            return x0/(x0*x0 + x1*x1)

        def ddlp_da1(tau):
            x0, x1 = A[0] + tau*B[0], A[1] + tau*B[1]
            r2 = x0*x0 + x1*x1
            # d/dx1 of -(n.x)/r2
            return -(N[1]*r2 - (N[0]*x0 + N[1]*x1)*2*x1)/r2**2

        res[0, i] = float(line_sum(slp, i))
        res[1, i] = float(line_sum(dlp, i))
        res[2, i] = float(line_sum(dslp_da0, i))
        res[3, i] = float(line_sum(ddlp_da1, i))
    np.savez(os.path.join(HERE, "qbx_line_taylor.npz"), a=a, b=b, n=n, order=order,
             values=res)


if __name__ == "__main__":
    p2p_fixture()
    derivative_fixture()
    qbx_fixture()
    print("wrote fixtures to", HERE)
