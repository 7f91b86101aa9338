"""Writes segment_story.json: the story's centrelines and clot, the input of the vessel geometry, taken from the
reference's geometry/segment.json (only the fields the geometry reads: mm per model unit, the junction, the three
branches' points, the clot's face point and body length). Run in a container that holds the reference:

    python tests/golden/make_segment.py /root/reference/geometry/segment.json
"""
import json
import os
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/geometry/segment.json"
with open(src) as fh:
    seg = json.load(fh)
out = {"mm_per_unit": float(seg["mm_per_unit"]),
       "junction_point": [float(x) for x in seg["junction_point"]],
       "branches": {name: {"points": [[float(x) for x in p] for p in seg["branches"][name]["points"]]}
                    for name in ("trunk", "blocked", "sibling")},
       "clot": {"face_point": [float(x) for x in seg["clot"]["face_point"]],
                "body_length": float(seg["clot"]["body_length"])}}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "segment_story.json"), "w") as fh:
    json.dump(out, fh, indent=0)
