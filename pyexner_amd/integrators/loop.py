"""Helpers shared by the integrators' output-window loops.

The dt of every step is produced on the device; the host reads it through a
side stream while the step it feeds is already running, so the GPU is not
left idle between steps.
"""
import numpy as np
import torch


class DtReader:
    """Copies a device scalar to pinned host memory on a side stream."""

    def __init__(self, dtype, device):
        self.host = torch.empty((), dtype=dtype, pin_memory=True)
        self.side = torch.cuda.Stream(device=device)
        self.ready = torch.cuda.Event()
        self.done = torch.cuda.Event()

    def start(self, dt_dev):
        self.ready.record(torch.cuda.current_stream())
        with torch.cuda.stream(self.side):
            self.side.wait_event(self.ready)
            self.host.copy_(dt_dev, non_blocking=True)
            dt_dev.record_stream(self.side)
            self.done.record(self.side)

    def value(self):
        self.done.synchronize()
        return float(self.host.item())


def next_target(time, out_freq, end_time):
    """forwardeuler.py:33-34: min((int(time/out_freq) + 1) * out_freq, end_time)."""
    return min((int(np.int32(time / out_freq)) + 1) * out_freq, end_time)
