"""Interior-cell assembly on the GPU (gdtp_hierarchy_create_ex with GDTP_ASSEMBLY_DEVICE) against
the host assembly of the same hierarchy: every deep stencil entry and every boundary-row entry of
mass and stiffness agrees to rounding (1e-12 of the array's largest entry), with identical sparsity.
The host assembly is itself pinned by closed forms and invariants (tests/test_assembly.py: textbook
Q1 stencils on flat cube faces, unit area, S 1 = 0, symmetry), so the device one is too; cases
cover flat faces, a curved seamed surface and an anisotropic per-triangle metric."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import gradientdomaintextureprocessing_amd as g  # noqa: E402


def _case(name, assembly):
    if name == "cube":
        mesh = g.TexturedMesh.cube(16, 128, 128)
        return g.InitializeHierarchy(mesh, 128, 128, 2, assembly=assembly)
    if name == "icosphere":
        mesh = g.TexturedMesh.icosphere(4, 256, 256, 0.05, 7)
        return g.InitializeHierarchy(mesh, 256, 256, 3, assembly=assembly)
    mesh = g.TexturedMesh.torus(64, 32, 4, 2, 256, 256, 1.0, 0.35, 0.05, 5)
    v, t, _ = mesh.arrays()
    vf = g.surface_gradient_field(v, t, np.sin(3 * v[:, 0]) + np.cos(2 * v[:, 1]) + v[:, 2])
    metric = g.vector_field_metric(g.embedding_metric(v, t), vf)
    return g.InitializeHierarchy(mesh, 256, 256, 3, triangle_metric=metric, assembly=assembly)


@pytest.mark.parametrize("name", ["cube", "icosphere", "anisotropic-torus"])
def test_device_assembly_matches_host(name):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    Hh, Hd = _case(name, "host"), _case(name, "device")
    assert Hh.level_info(0) == Hd.level_info(0)
    worst = 0.0
    for op in ("mass", "stiffness"):
        a, b = Hh.array(0, f"{op}.deep", np.float64), Hd.array(0, f"{op}.deep", np.float64)
        scale = np.abs(a).max()
        worst = max(worst, np.abs(a - b).max() / scale)
        assert np.abs(a - b).max() <= 1e-12 * scale, op
        for part in ("bb", "bd"):
            for idx in ("rowptr", "col"):
                assert np.array_equal(Hh.array(0, f"{op}.{part}.{idx}"), Hd.array(0, f"{op}.{part}.{idx}")), (op, part, idx)
            a, b = Hh.array(0, f"{op}.{part}.val", np.float64), Hd.array(0, f"{op}.{part}.val", np.float64)
            assert np.abs(a - b).max() <= 1e-12 * max(np.abs(a).max(), scale), (op, part)
    print(f"[assembly {name}] worst deep-stencil difference {worst:.1e} of the largest entry", flush=True)
