"""Writes ref_line_counts.json: the number of lines of every source and
header file of the reference (read from /root/reference), so that the CPU
test of citations can check every `src/...:N-M` reference in this repository
without the reference being present."""
import json
import os

REF = "/root/reference"
out = {}
for sub in ("src", "include"):
    for name in sorted(os.listdir(os.path.join(REF, sub))):
        path = os.path.join(REF, sub, name)
        if os.path.isfile(path):
            with open(path, errors="replace") as f:
                out[f"{sub}/{name}"] = sum(1 for _ in f)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_line_counts.json"), "w") as f:
    json.dump(out, f, indent=1, sort_keys=True)
    f.write("\n")
print(len(out), "files")
