"""The Rys table's per-n block sizes hard-coded for the LDS staging in
tonto_amd/csrc/tgpu_device.h (rys_block_size) match the generated table."""
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tonto_amd", "csrc")


def test_block_sizes_match_table():
    inc = open(os.path.join(CSRC, "rys_table.inc")).read()
    offs = [int(x) for x in re.search(r"rys_table_offset\[RYS_NMAX \+ 1\] = \{([^}]*)\}", inc).group(1).split(",")]
    total = int(re.search(r"#define RYS_TABLE_SIZE (\d+)", inc).group(1))
    ncheb = int(re.search(r"#define RYS_NCHEB (\d+)", inc).group(1))
    dev = open(os.path.join(CSRC, "tgpu_device.h")).read()
    body = re.search(r"constexpr int rys_block_size\(int n\)\s*\{(.*?)\}", dev, re.S).group(1)
    nint = [int(x) for x in re.findall(r"\? (\d+)", body)] + [int(re.search(r": (\d+)\)", body).group(1))]
    sizes = [offs[n + 1] - offs[n] for n in range(1, len(offs) - 1)] + [total - offs[-1]]
    for n, (ni, sz) in enumerate(zip(nint, sizes), start=1):
        assert ni * ncheb * 2 * n == sz, (n, ni, sz)
        assert offs[n] % 2 == 0          # 16-byte aligned double2 reads
