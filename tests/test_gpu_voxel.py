"""GPU: per-vertex material ids computed on the device (fame_plan_set_structure) against the
host restatement (structures.woodpile_sc_ids / sphere_ids / tetra_ids, itself checked in
tests/test_abi_cpu.py), bit for bit on the benchmark structures at their real sizes, and a
plan voxelised on the device gives the host-voxelised plan's fncB bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


CASES = [
    ("woodpile", (24, 24, 24), 0.25), ("woodpile", (60, 60, 60), 0.25), ("woodpile", (6, 10, 12), 0.25),
    ("sphere", (48, 48, 48), 0.12), ("sphere", (72, 72, 72), 0.12), ("sphere", (8, 8, 8), 0.3),
    ("tetra", (60, 60, 60), 0.3), ("tetra", (8, 8, 8), 0.6), ("sphere", (10, 12, 9), 0.2),
]


@pytest.mark.parametrize("kind,dims,size", CASES)
def test_device_voxelisation_matches_host(pkg, torch_mod, kind, dims, size):
    print(f"device voxelisation {kind} dims={dims} size={size}", flush=True)
    S = pkg.structures
    A = S.SC_LATTICE if kind == "woodpile" else S.FCC_LATTICE
    host = {"woodpile": lambda: S.woodpile_sc_ids(dims, width=size),
            "sphere": lambda: S.sphere_ids(dims, size, A),
            "tetra": lambda: S.tetra_ids(dims, size, A)}[kind]()
    table = S.bi_isotropic_table(13.0, 0.2, A)
    theta = S.theta_of_k(np.array([0.3, 0.2, 0.1]), A)
    plan = pkg.Plan(dims, theta, table, None)
    plan.set_structure(kind, size, lattice=A)
    dev = plan.material_ids().cpu().numpy()
    assert 0 < host.sum() < host.size
    assert np.array_equal(dev, host), int((dev != host).sum())
    plan.close()


def test_device_voxelised_plan_fncB_bitwise(pkg, torch_mod):
    print("fncB: device-voxelised plan == host-voxelised plan", flush=True)
    S = pkg.structures
    dims = (24, 24, 24)
    A = S.FCC_LATTICE
    thetas = np.stack([S.theta_of_k(np.array(k), A) for k in ([0.3, 0.2, 0.1], [1.0, 0.5, 0.0])])
    table = S.bi_isotropic_table(13.0, 2.72, A)
    ph = pkg.Plan(dims, thetas, table, S.sphere_ids(dims, 0.12, A))
    pd = pkg.Plan(dims, thetas, table, None)
    pd.set_structure("sphere", 0.12, lattice=A)
    rng = np.random.default_rng(3)
    x = torch_mod.as_tensor(rng.standard_normal((2, 4 * ph.N)) + 1j * rng.standard_normal((2, 4 * ph.N)))
    assert torch_mod.equal(ph.fncB(x), pd.fncB(x))
    ph.close()
    pd.close()
