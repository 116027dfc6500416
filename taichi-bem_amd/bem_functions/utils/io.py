"""(N, ..., n) real <-> complex views (src/BEM_functions/utils/io.py:4-31)."""
import torch


def warp_tensor(x: torch.Tensor) -> torch.Tensor:
    n_ = x.shape[-1]
    if n_ == 1:
        return x.squeeze(-1)
    if n_ == 2:
        return torch.view_as_complex(x.contiguous())
    raise NotImplementedError(
        f"last dimension {n_} has no complex reading: it must be 1 (real values) or 2 (real, imaginary pairs)"
    )


def unwarp_tensor(x: torch.Tensor) -> torch.Tensor:
    if x.is_complex():
        return torch.view_as_real(x)
    return x.unsqueeze(-1)
