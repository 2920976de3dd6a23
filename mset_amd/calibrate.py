"""Known-size copy for calibrating rocprofv3 memory counters.

    rocprofv3 --pmc FETCH_SIZE -- python -m mset_amd.calibrate [MiB]

Copies a device buffer of the given size (default 1024 MiB) once with a plain copy kernel, so that the
FETCH_SIZE / WRITE_SIZE a profiler reports for that dispatch can be compared with the known byte count
(reads = writes = size) and the same factor applied to the multislice kernels' counters.
"""
import sys


def main():
    import torch
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print('copied %d MiB (%d bytes read, %d bytes written)' % (mib, n * 4, n * 4))


if __name__ == '__main__':
    main()
