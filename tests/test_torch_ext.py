"""CPU checks of the LibTorch extension: it was built, imports against the
HIP library, and refuses CPU tensors (there is no CPU path)."""
import pytest

torch = pytest.importorskip("torch")


def test_extension_imports_and_rejects_cpu_tensors():
    from magnetohydrodynamics_simulator_amd import torch_ops

    ext = torch_ops.load()
    assert hasattr(ext, "imex_steps")
    f = {k: torch.zeros(8, 8, dtype=torch.float64) for k in torch_ops.FIELDS}
    with pytest.raises(RuntimeError, match="GPU tensors"):
        torch_ops.imex_steps(f, 1e-3, 0.1, 0.1, 1e-3, 1)
