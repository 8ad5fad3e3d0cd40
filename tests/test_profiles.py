"""The committed rocprofv3 PMC summaries that bench.py reads for its roofline
`traffic` field: each parses, its traffic line equals the total of the step
kernel's row in the same table, and the traffic lies between the kernel's
algorithmic bytes (2 Q s per node, D3Q19) and 1.25 times them."""
import os
import re

import pytest

from bench import pmc_traffic

NODES = 256 ** 3


@pytest.mark.parametrize("storage,elem,row", [
    ("fp64", 8, r"\| k_step<3,double,1> \(fixed markers\)"),
    ("fp16", 2, r"\| k_step_lds<3,_Float16,1> \(fixed markers, pair sums inside\)"),
])
def test_pmc_summary_traffic_is_the_step_rows_total(storage, elem, row):
    t = pmc_traffic(storage, 256)
    assert t is not None
    algorithmic = 2 * 19 * elem * NODES
    assert algorithmic <= t <= 1.25 * algorithmic
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "profiles",
                        f"pmc_k_step_d3q19_256cube_{storage}.md")
    with open(path) as fh:
        text = fh.read()
    line = next(ln for ln in text.splitlines() if re.match(row, ln))
    total_gb = float(re.findall(r"([0-9.]+) GB", line)[2])   # read, written, total
    assert abs(total_gb * 1e9 - t) <= 0.5e6
