"""Write the golden fixtures of the reference's own tests into tests/golden/.

Reads, from /root/reference (the reference checkout, present only where the fixtures
are regenerated):
  Core/tests/grid_input.roi            the 10-point ROI of the DataCloud test
  Core/tests/test_DataCloud.cpp:129-139  its hard-coded 10x10 neighbour table
and writes datacloud_10nn.json = {"points": [[x,y,z]...], "labels": [...],
"neighbours": [[...10 ints...] x 10]}.  Data only: coordinates and indices.
"""
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent


def main():
    labels, pts = [], []
    for line in (REF / "Core/tests/grid_input.roi").read_text().splitlines():
        f = line.split()
        if len(f) == 4:
            labels.append(int(f[0]))
            pts.append([float(v) for v in f[1:]])
    src = (REF / "Core/tests/test_DataCloud.cpp").read_text()
    block = re.search(r"std::vector<int> row = \{(.*?)\};", src, re.S).group(1)
    ints = [int(v) for v in re.findall(r"-?\d+", block)]
    assert len(ints) == 100, len(ints)
    table = [ints[i * 10:(i + 1) * 10] for i in range(10)]
    (OUT / "datacloud_10nn.json").write_text(json.dumps({"labels": labels, "points": pts, "neighbours": table},
                                                        indent=1))
    print("wrote", OUT / "datacloud_10nn.json")


if __name__ == "__main__":
    main()
