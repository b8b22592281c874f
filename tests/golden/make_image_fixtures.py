"""Writes the JPEG decode fixtures under tests/golden/images/.

Each fixture is a small JPEG file written by Pillow and the RGB pixels that
Pillow's libjpeg-turbo decodes from it with its defaults (accurate integer
IDCT, triangle-filter chroma upsampling): the same decode the reference gets
from tjDecompress2(TJPF_RGBA, TJFLAG_ACCURATEDCT) (src/core/utility/image.c:194-279).

    python tests/golden/make_image_fixtures.py
"""

import io
import os

import numpy as np
from PIL import Image

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "images")

CASES = [
    # name, width, height, mode, subsampling, progressive, quality, extra save options
    ("rgb_420_progressive", 45, 29, "RGB", 2, True, 85, {}),
    ("rgb_422_restart", 38, 21, "RGB", 1, False, 70, {"restart_marker_blocks": 2}),
    ("rgb_444", 17, 13, "RGB", 0, False, 95, {}),
    ("grey_progressive", 23, 31, "L", 0, True, 60, {}),
]


def source(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([np.sin(x / 4.0) * 100 + 128, np.cos(y / 3.0) * 90 + 128, (x * 7 + y * 11) % 256], -1)
    return np.clip(img + rng.normal(0, 25, (h, w, 3)), 0, 255).astype(np.uint8)


def main():
    os.makedirs(HERE, exist_ok=True)
    for i, (name, w, h, mode, sub, prog, q, extra) in enumerate(CASES):
        img = source(w, h, i)
        im = Image.fromarray(img if mode == "RGB" else img[..., 0], mode)
        bio = io.BytesIO()
        im.save(bio, "JPEG", quality=q, subsampling=sub, progressive=prog, **extra)
        data = bio.getvalue()
        rgb = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
        with open(os.path.join(HERE, name + ".jpg"), "wb") as f:
            f.write(data)
        np.save(os.path.join(HERE, name + "_rgb.npy"), rgb)
        print(name, len(data), "bytes")


if __name__ == "__main__":
    main()
