"""Writes tests/golden/veach_ajar_sty_materials.json: the "materials" object (every material's
shader graph, unchanged) of the reference's stylized scene scenes/veach-ajar-sty-v2/scene-sty.json.
The fixture is data only. Run in a checkout that has the reference at /root/reference."""
import json
import os

SRC = "/root/reference/scenes/veach-ajar-sty-v2/scene-sty.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "veach_ajar_sty_materials.json")

if __name__ == "__main__":
    with open(SRC) as f:
        scene = json.load(f)
    with open(OUT, "w") as f:
        json.dump({"materials": scene["materials"]}, f, indent=0, sort_keys=True)
    print(OUT, os.path.getsize(OUT))
