"""Plain device copy of a known size, for scaling rocprofv3's FETCH_SIZE and
WRITE_SIZE counters (KiB per dispatch) to bytes: run under
`rocprofv3 --pmc FETCH_SIZE` (and WRITE_SIZE) and divide the counter by
the copy's size.  Prints the size it copied."""
import sys

import torch

MIB = int(sys.argv[1]) if len(sys.argv) > 1 else 512
n = MIB * 1024 * 1024 // 8
x = torch.rand(n, dtype=torch.float64, device="cuda")
y = torch.empty_like(x)
for _ in range(5):
    y.copy_(x)
torch.cuda.synchronize()
print(f"copied {MIB} MiB x5 (float64, {n} elements)", flush=True)
