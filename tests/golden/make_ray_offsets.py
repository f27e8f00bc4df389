"""Write ray_offsets.json: the 256-entry ray-offset (jitter) table that
RobartsVTK's casters index by (x mod 16) + 16 (y mod 16).

Source: Visualization/vtkCudaVisualization/vtkCudaVolumeMapper.cxx:59-93
(RobartsVTK, MIT licence).  This script reads that file as text in the
container that holds the reference and writes only the numbers.
"""
import json
import re
import sys
from pathlib import Path

SRC = Path("/root/reference/Visualization/vtkCudaVisualization/vtkCudaVolumeMapper.cxx")


def main() -> None:
    text = SRC.read_text()
    m = re.search(r"RANDOM_RAY_OFFSETS\[256\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        sys.exit("table not found")
    vals = [v.strip().rstrip("f") for v in m.group(1).split(",") if v.strip()]
    assert len(vals) == 256, len(vals)
    out = {"source": "Visualization/vtkCudaVisualization/vtkCudaVolumeMapper.cxx:59-93",
           "licence": "MIT (RobartsVTK LICENSE.md)",
           "index": "(x % 16) + 16 * (y % 16)",
           "values": vals}
    Path(__file__).with_name("ray_offsets.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
