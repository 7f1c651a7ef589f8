"""HIP graph capture of a host function's launches.

`capture(body)` records every launch `body()` makes into a new HIP graph on
a dedicated capture stream and returns the graph without running it; the
caller replays it. Unlike `torch.cuda.graph` it does not synchronise the
device, run the Python garbage collector or empty torch's allocator cache
around the capture, which made a first Newton step cost several times a
steady one. Tensors `body` allocates come from the graph's private pool.
"""

import torch

_STREAMS = {}


def capture(body, device=None):
    dev = torch.device(device) if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    key = dev.index if dev.index is not None else torch.cuda.current_device()
    s = _STREAMS.get(key)
    if s is None:
        s = _STREAMS[key] = torch.cuda.Stream(dev)
    cur = torch.cuda.current_stream(dev)
    s.wait_stream(cur)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        g.capture_begin()
        try:
            body()
        finally:
            g.capture_end()
    cur.wait_stream(s)
    return g
