"""Writes the seed files of the image-loader fuzz run (tests/sanitize): small
baseline JPEGs covering 4:4:4, 4:2:0 with restart intervals, and grayscale.
Run from this directory: python gen_fuzz_seeds.py"""
import numpy as np
from PIL import Image


def image(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w] / max(h, w)
    img = np.stack([127 + 100 * np.sin(4 * x + c) * np.cos(3 * y - c) for c in (0.0, 1.0, 2.0)], -1)
    return np.clip(img + rng.normal(0, 10, img.shape), 0, 255).astype(np.uint8)


if __name__ == "__main__":
    Image.fromarray(image(19, 27, 1)).save("fuzz_seed_444.jpg", "JPEG", quality=85, subsampling=0)
    Image.fromarray(image(24, 40, 2)).save("fuzz_seed_420_rst.jpg", "JPEG", quality=80, subsampling=2,
                                           restart_marker_blocks=2)
    Image.fromarray(image(16, 16, 3)[..., 0], "L").save("fuzz_seed_gray.jpg", "JPEG", quality=90)
