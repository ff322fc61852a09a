"""Host <-> device helpers for the GPU tests (torch as device memory)."""
import numpy as np
import torch

import tmv_amd


def to_dev(a, dtype=torch.float64, layout="col"):
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64)))
    t = t.to(device="cuda", dtype=dtype)
    if t.dim() == 2 and layout == "col":
        t = tmv_amd.colmajor(t)
    return t


def to_host(t):
    return t.detach().cpu().double().numpy()
