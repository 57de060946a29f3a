"""cyx_chebyshev_filter_noalloc against cyx_chebyshev_filter, through the
C-ABI: the same kernels with the KB scratch taken from the library's device
array instead of allocated buffers, so the results must be bitwise equal.
Two calls in a row (the second with fewer columns, so it reuses the array
with stale contents from the first), on a helical Bloch channel with
projectors (complex128 and complex64) and a cyclic Gamma channel (FP64);
and a call whose scratch would exceed the array is refused before anything
runs."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CASES = {
    "helical_bloch": dict(nr=30, nt=16, nz=14, twist=0.05, kpt=(0.0, 1.0, 0.6), dtypes=("c128", "c64")),
    "cyclic_gamma": dict(nr=24, nt=16, nz=14, twist=0.0, kpt=(0.0, 0.0, 0.0), dtypes=("f64",)),
}
TORCH = {"f64": torch.float64, "c128": torch.complex128, "c64": torch.complex64}


def _filter(ham, fn, X, m, a, b, a0):
    from cyclix_msparc_amd import _native as nat
    from cyclix_msparc_amd.api import _dtype_code
    Xw = X.clone()
    Y = torch.empty_like(X)
    rc = fn(ham.h, _dtype_code(X), C.c_void_p(Xw.data_ptr()), C.c_void_p(Y.data_ptr()), X.shape[0], ham.N, m,
            float(a), float(b), float(a0))
    nat.check(rc)
    torch.cuda.synchronize()
    return Y


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_filter_matches_filter(device, name):
    from cyclix_msparc_amd.api import Hamiltonian
    from cyclix_msparc_amd.system import small_system
    c = CASES[name]
    sysm = small_system(nr=c["nr"], nt=c["nt"], nz=c["nz"], twist=c["twist"], kpt=c["kpt"], nstates=8, npl=8)
    g = sysm.grid
    r = np.tile(g.r, g.nt * g.nz)
    veff = -1.5 * np.exp(-((r - sysm.atoms[:, 0].max()) / 1.2) ** 2)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, sysm.projectors())
    assert ham._proj is not None
    lib = device.lib
    rng = np.random.default_rng(3)
    for dt in c["dtypes"]:
        for ncol in (11, 4):  # the second call is smaller
            X = rng.standard_normal((ncol, g.N))
            if dt != "f64":
                X = X + 1j * rng.standard_normal((ncol, g.N))
            Xd = torch.from_numpy(X).to(device.torch_device).to(TORCH[dt]).contiguous()
            print(f"noalloc filter {name} {dt} ncol={ncol}", flush=True)
            ref = _filter(ham, lib.cyx_chebyshev_filter, Xd, 8, 0.3, 40.0, -2.0)
            got = _filter(ham, lib.cyx_chebyshev_filter_noalloc, Xd, 8, 0.3, 40.0, -2.0)
            assert torch.equal(got, ref), (name, dt, ncol)
    ham.close()


def test_noalloc_filter_refuses_oversized_scratch(device):
    """The scratch bound is checked on the host before anything is enqueued:
    a column count whose KB scratch exceeds 16 MiB returns CYX_ERR_ARG, and
    the buffers (large enough for one column only) are never touched."""
    from cyclix_msparc_amd import _native as nat
    from cyclix_msparc_amd.api import Hamiltonian
    from cyclix_msparc_amd.system import small_system
    sysm = small_system(nr=30, nt=16, nz=14, twist=0.05, kpt=(0.0, 1.0, 0.6), nstates=8, npl=8)
    g = sysm.grid
    ham = Hamiltonian(device, sysm, sysm.kpt, np.zeros(g.N), sysm.projectors())
    X = torch.zeros((1, g.N), dtype=torch.complex128, device=device.torch_device)
    Y = torch.zeros_like(X)
    ncol = 1 << 24
    rc = device.lib.cyx_chebyshev_filter_noalloc(ham.h, nat.CYX_C128, C.c_void_p(X.data_ptr()),
                                                 C.c_void_p(Y.data_ptr()), ncol, g.N, 8, 0.3, 40.0, -2.0)
    assert rc == nat.CYX_ERR_ARG
    msg = device.lib.cyx_last_error().decode()
    assert "maximum" in msg, msg
    torch.cuda.synchronize()
    assert torch.count_nonzero(Y).item() == 0
    ham.close()
