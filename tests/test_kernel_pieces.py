"""One try's pieces from the HIP kernels' own device functions (tests/kernel_probe.hip), against numpy and the oracle:
the wave reductions, F and the exact Jacobian dF/dN = Jn + w 1^T, the factorization's solve of M0, and the four
Rodas3 stages against a dense numpy solve of W (GPU)."""

from pathlib import Path

import numpy as np
import oracle
import pytest
import torch

from maxichem_amd import _hip, _kernels, _tables, standin

pytestmark = pytest.mark.gpu
PROBE = Path(__file__).resolve().parent / "kernel_probe.hip"


def probe(net, photo, x, T, P, h, J):
    tab = _tables.build(net, photo)
    src = _kernels.source(tab, 1, 1e-6) + "\n" + PROBE.read_text()
    mod = _hip.Module(_hip.compile_source(src, "probe.hip"), torch.cuda.current_device())
    fn = mod.function("mxc_probe")
    ns = tab.ns
    perm = tab.perm

    def dev(a, dt=torch.float64):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt, device="cuda")

    host = _kernels.host_tables(tab)  # the solver's own tables, in the kernels' Net field order
    tabs = [torch.from_numpy(host[k]).cuda() for k in _kernels.NET_FIELDS]
    net_struct = torch.tensor([a.data_ptr() for a in tabs], dtype=torch.int64, device="cuda")
    out = torch.zeros(8 + 8 * ns + 2 * ns * ns, dtype=torch.float64, device="cuda")
    xs = dev(x)
    Jt = dev(J if J is not None else np.zeros(1))
    import ctypes

    args = [_hip.ptr(net_struct), _hip.ptr(xs), ctypes.c_double(T), ctypes.c_double(P), ctypes.c_double(h),
            _hip.ptr(Jt), _hip.ptr(out)]  # fmt: skip
    _hip.launch(fn, 1, 64, 8 * tab.layout()["SIZE"], torch.cuda.current_stream().cuda_stream, args)
    torch.cuda.synchronize()
    mod.unload()
    o = out.cpu().numpy()
    inv = np.empty(ns, np.int64)
    inv[perm] = np.arange(ns)

    def net_order(v):  # elimination order -> network order
        return v[inv]

    b = 8
    got = {"sums": o[:4], "F": net_order(o[b : b + ns]), "w": net_order(o[b + ns : b + 2 * ns])}
    jn = o[b + 2 * ns : b + 2 * ns + ns * ns].reshape(ns, ns)
    got["mjn"] = jn[np.ix_(inv, inv)]
    got["lu"] = o[b + 2 * ns + ns * ns : b + 2 * ns + 2 * ns * ns].reshape(ns, ns)  # elimination order
    got["dinv"] = o[b + 2 * ns + 2 * ns * ns : b + 3 * ns + 2 * ns * ns]
    got["perm"] = perm
    got["k0"] = tab.k0
    v = b + 3 * ns + 2 * ns * ns
    for k, name in enumerate(("m0f", "K1", "K2", "K3", "K4")):
        got[name] = net_order(o[v + k * ns : v + (k + 1) * ns])
    return tab, got


CASES = [("synthetic_small", False), ("synthetic", True), ("synthetic_decay", True), ("NCHO", True)]


@pytest.mark.parametrize("stem,photo", CASES)
def test_one_try_matches_numpy(read_network, stem, photo):
    net = standin.load(stem) if stem == "NCHO" else read_network(stem)
    rng = np.random.default_rng(4)
    x = 10.0 ** rng.uniform(-8, 0, len(net.species))
    x /= x.sum()
    T, P, h = 1700.0, 3e5, 1e-4
    J = np.full(len(net.photo), 1e-3) if photo else None
    tab, got = probe(net, photo, x, T, P, h, J)
    assert got["sums"][0] == pytest.approx(sum(i + 0.5 for i in range(64)), rel=0, abs=0)
    assert got["sums"][1] == pytest.approx(sum(2.0**-i for i in range(64)), rel=1e-15)
    assert got["sums"][2] == 63.0 and got["sums"][3] == 17 * 3.25
    orc = oracle.Oracle(tab)
    F, JN = orc.rhs_jacobian(T, P, x, J)
    scale = np.abs(F).max()
    np.testing.assert_allclose(got["F"], F, rtol=1e-10, atol=1e-12 * scale)
    # the exact Jacobian: Jn + w 1^T, with the kernel's -Jn and w
    JK = -got["mjn"] + got["w"][:, None]
    np.testing.assert_allclose(JK, JN, rtol=1e-9, atol=1e-12 * np.abs(JN).max())
    W0 = np.eye(len(x)) / (0.5 * h) + got["mjn"]  # M0 = I / (gamma h) - Jn
    # the factors: LU without row exchanges of M0 in the elimination order
    A = W0[np.ix_(got["perm"], got["perm"])].copy()
    n = len(A)
    for k in range(n):
        A[k + 1 :, k] /= A[k, k]
        A[k + 1 :, k + 1 :] -= np.outer(A[k + 1 :, k], A[k, k + 1 :])
    k0 = int(got["k0"])
    U = A.copy()
    for k in range(k0, n):  # the trailing block keeps its U columns scaled: U_ik / U_kk above the diagonal
        A[k0:k, k] /= U[k, k]
    bad = np.argwhere(~np.isclose(got["lu"], A, rtol=1e-9, atol=1e-12 * np.abs(A).max()))
    assert len(bad) == 0, f"LU factors differ at (row, column) {[tuple(map(int, q)) for q in bad[:40]]}, " + str(
        [(float(got["lu"][i, j]), float(A[i, j])) for i, j in bad[:8]])
    np.testing.assert_allclose(got["dinv"], 1.0 / np.diag(U), rtol=1e-12, err_msg="1 / U_ii")
    m0f = np.linalg.solve(W0, F)
    np.testing.assert_allclose(got["m0f"], m0f, rtol=1e-8, atol=1e-10 * np.abs(m0f).max())
    W = np.eye(len(x)) / (0.5 * h) - JN
    k1 = np.linalg.solve(W, F)
    k2 = np.linalg.solve(W, F + 4.0 * k1 / h)
    f3 = orc.rhs_jacobian(T, P, x + 2 * k1, J)[0]
    k3 = np.linalg.solve(W, f3 + (k1 - k2) / h)
    f4 = orc.rhs_jacobian(T, P, x + 2 * k1 + k3, J)[0]
    k4 = np.linalg.solve(W, f4 + (k1 - k2 - 8.0 / 3.0 * k3) / h)
    for name, want in (("K1", k1), ("K2", k2), ("K3", k3), ("K4", k4)):
        # the stages are increments of a state of order 1: rounding at 1e-13 of it is the floor
        np.testing.assert_allclose(got[name], want, rtol=1e-7, atol=1e-13 * np.abs(x).max(), err_msg=name)
