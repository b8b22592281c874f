"""Writes vkrt_amd/csrc/sheen_ltc_table.inc from the reference's sheen LTC data.

The table is Blender Cycles' fit of the Zeltner-Burley-Chiang 2022 microfiber
sheen (intern/cycles/scene/shader.tables), which the reference embeds in
src/shaders/bsdf/data/sheen_ltc.slang: three 32 x 32 layers (LTC transform
a, transform b, directional albedo) indexed [layer][roughness][cos theta].
It is data both the device lobe and the oracle read; this script was run once
and its output is committed.

    python vkrt_amd/csrc/tools/extract_sheen_ltc.py /root/reference/src/shaders/bsdf/data/sheen_ltc.slang
"""

import os
import re
import sys


def main(src):
    text = open(src).read()
    body = text[text.index("SHEEN_LTC_TABLE[3072]"):]
    body = body[body.index("{") + 1:body.index("}")]
    values = [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", body)]
    assert len(values) == 3072, len(values)
    out = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sheen_ltc_table.inc")
    with open(out, "w") as f:
        f.write("/* Microfiber sheen LTC fit (Zeltner, Burley, Chiang 2022) as tabulated by Blender\n"
                " * Cycles: layers transform-a, transform-b, albedo; 32 roughness rows of 32\n"
                " * cos-theta columns each. Written by tools/extract_sheen_ltc.py. */\n")
        for layer in range(3):
            f.write(f"/* layer {layer} */\n")
            for row in range(32):
                vals = values[layer * 1024 + row * 32: layer * 1024 + row * 32 + 32]
                f.write(", ".join(f"{v:.5f}f" for v in vals) + ",\n")


if __name__ == "__main__":
    main(sys.argv[1])
