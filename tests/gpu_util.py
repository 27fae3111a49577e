"""Helpers shared by the GPU parity tests."""
import numpy as np


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    if den == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / den)


def to_dev_u32(torch, arr):
    a = np.ascontiguousarray(arr, np.uint32)
    return torch.from_numpy(a.view(np.int32).copy()).cuda()


def from_dev_u32(t):
    return t.cpu().numpy().view(np.uint32)


def say(msg: str) -> None:
    print(msg, flush=True)
