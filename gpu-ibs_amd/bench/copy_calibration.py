"""A plain 2 GiB device-to-device copy, three times: the known-size
reference that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters
(measurement only)."""
import torch

x = torch.ones(1 << 28, dtype=torch.int64, device="cuda")  # 2 GiB
y = torch.empty_like(x)
torch.cuda.synchronize()
for _ in range(3):
    y.copy_(x)
torch.cuda.synchronize()
print("copied 3 x 2 GiB", flush=True)
