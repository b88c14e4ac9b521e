"""Device selection, mirroring pydiffvg/device.py:1-25.

The render path of this package runs on the GPU only; ``set_use_gpu(False)`` is
accepted for API compatibility, but rendering then raises, because there is no CPU
fallback in the product path.
"""
import torch

use_gpu = torch.cuda.is_available()
device = torch.device('cuda') if use_gpu else torch.device('cpu')


def set_use_gpu(v):
    global use_gpu
    global device
    use_gpu = v
    if not use_gpu:
        device = torch.device('cpu')
    elif device.type != 'cuda':
        device = torch.device('cuda')


def get_use_gpu():
    return use_gpu


def set_device(d):
    global device
    global use_gpu
    device = d
    use_gpu = device.type == 'cuda'


def get_device():
    return device
