"""(N, ..., n) real <-> complex views (src/BEM_functions/utils/io.py:4-31)."""
import torch


def warp_tensor(x: torch.Tensor) -> torch.Tensor:
    n_ = x.shape[-1]
    if n_ == 1:
        return x.squeeze(-1)
    if n_ == 2:
        return torch.view_as_complex(x.contiguous())
    raise NotImplementedError(
        "A Tensor with shape = (*, {}) can not be warped, "
        "we only support 1 (Scalar) or 2 (Complex) for the last dim".format(n_)
    )


def unwarp_tensor(x: torch.Tensor) -> torch.Tensor:
    if x.is_complex():
        return torch.view_as_real(x)
    return x.unsqueeze(-1)
