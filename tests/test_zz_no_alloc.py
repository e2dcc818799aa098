"""salmon_hip_density_noalloc (partial sums in the library's per-GPU device array, no
allocation in the call) against salmon_hip_density through the C ABI.

Two calls in a row on the same GPU, the second on a smaller system, each equal to the
allocating entry point bit for bit (same kernels, same summation order); then a system whose
partial sums exceed the array is rejected on the host with SALMON_HIP_EINVAL.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def _density_pair(torch, s):
    from salmon2_amd.hip import Handle

    h = Handle(s, stream=torch.cuda.current_stream())
    psi = torch.from_numpy(s.psi).cuda()
    rho_a = torch.zeros((s.nspin, s.nz, s.ny, s.nx), dtype=torch.float64, device="cuda")
    rho_b = torch.full_like(rho_a, float("nan"))
    h.density(psi, rho_a)
    h.density_noalloc(psi, rho_b)
    torch.cuda.synchronize()
    h.close()
    return rho_a.cpu().numpy(), rho_b.cpu().numpy()


def test_density_noalloc_two_calls_equal_allocating_entry(torch_dev):
    from salmon2_amd import system

    torch = torch_dev
    # 32^3, 16 orbitals x 8 k = 4 partial groups; then 16^3, 40 orbitals x 1 k = 2 groups
    for name, kw in (("c2", dict(no=16, nt=2)), ("c1", dict(no=40, nt=2))):
        s = system.config(name, **kw)
        print(f"[gpu] density_noalloc {name} {s.nx}^3 x {s.no * s.nk} orbitals", flush=True)
        a, b = _density_pair(torch, s)
        print(f"[parity] density_noalloc {name}: max |diff| {np.abs(a - b).max():.3e}", flush=True)
        assert np.array_equal(a, b)
        # the density integrates to the electron count (normalised orbitals, rocc 2, weights 1/nk)
        assert abs(a.sum() * s.hvol - 2.0 * s.no) <= 1e-9 * s.no


def test_density_noalloc_rejects_beyond_scratch(torch_dev):
    from salmon2_amd import system
    from salmon2_amd.hip import Handle, SalmonHipError

    torch = torch_dev
    # 64^3 points x 33 groups of 32 orbitals = 8.65 M partial sums > the 8 Mi doubles
    s = system.make_si_system(n=64, no=1056, nt=2, with_projectors=False, host_orbitals=False, name="big")
    print("[gpu] density_noalloc rejection 64^3 x 1056 orbitals", flush=True)
    h = Handle(s, stream=torch.cuda.current_stream())
    psi = torch.zeros((s.nk, s.no, s.nspin, s.nz, s.ly, s.nx), dtype=torch.complex128, device="cuda")
    rho = torch.zeros((s.nspin, s.nz, s.ny, s.nx), dtype=torch.float64, device="cuda")
    with pytest.raises(SalmonHipError, match="exceed the device scratch"):
        h.density_noalloc(psi, rho)
    h.close()
    del psi
    torch.cuda.empty_cache()
