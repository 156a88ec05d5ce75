"""Writes tests/golden/enums.json: the ordinals of the reference's KernelName and
BufferName enums (KernelName.cs:11-21, BufferName.cs:11-63), read as text from
the reference tree.  Run in the build container only; the JSON is the fixture."""
import json
import re
from pathlib import Path

REF = Path("/root/reference/SvoTracer/SvoTracer.Kernel/Enums")


def ordinals(path: Path, enum: str):
    text = path.read_text(encoding="utf-8-sig")
    body = re.search(r"enum\s+" + enum + r"\s*\{(.*?)\}", text, re.S).group(1)
    names = [n.strip() for n in re.sub(r"//.*", "", body).replace("\n", " ").split(",") if n.strip()]
    return {n: i for i, n in enumerate(names)}


if __name__ == "__main__":
    out = {
        "KernelName": ordinals(REF / "KernelName.cs", "KernelName"),
        "BufferName": ordinals(REF / "BufferName.cs", "BufferName"),
    }
    dst = Path(__file__).with_name("enums.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst}: {len(out['KernelName'])} kernels, {len(out['BufferName'])} buffers")
