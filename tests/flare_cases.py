"""Inputs of the flare checks (our own data): seeded test images and the
parameter sets, shared by the oracle tests, the GPU tests and the fixture
generator."""
import numpy as np

from marmel_amd import abi


def params(enabled=1, intensity=0.6, threshold=0.85, streak=0.25, tint=(0.15, 0.45, 1.0)):
    return abi.FlareParams(enabled=enabled, intensity=intensity, threshold=threshold, streak_length=streak,
                           tint=abi.Vec3(*tint))


# name -> FlareParams: the defaults, and every clamp of the reference
PARAMS = {
    "default": params(),
    "strong": params(intensity=2.5, threshold=0.5, streak=0.8, tint=(1.0, 0.5, 0.2)),
    "thresh_low": params(threshold=-1.0, intensity=0.3),          # clamps to 0
    "thresh_high": params(threshold=1.5, intensity=8.0),          # clamps to 0.999
    "short": params(streak=0.001, intensity=3.0, threshold=0.6),  # clamps to 0.005
    "long": params(streak=3.0, threshold=0.7),                    # clamps to 1
    "no_tint": params(tint=(0.0, 0.0, 0.0), threshold=0.7),        # default tint
}

# these leave the image as it is
NOOP = {"disabled": params(enabled=0), "zero": params(intensity=0.0), "negative": params(intensity=-1.0)}


def reference_pattern():
    """The reference's own test image (tests/test_anamorphic.c:144-158):
    64 x 16 grey 100 with one white pixel at row 8, column 32."""
    img = np.full((16, 64, 3), 100, np.uint8)
    img[8, 32] = 255
    return img


def speckle(w, h, seed):
    """Seeded random image with saturated speckles, bright bars and dark rows."""
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 200, (h, w, 3), dtype=np.uint8)
    n = max(1, w * h // 40)
    ys, xs = rng.integers(0, h, n), rng.integers(0, w, n)
    img[ys, xs] = rng.integers(200, 256, (n, 3), dtype=np.uint8)
    img[h // 3, :] = 255
    img[:, w // 2] = [255, 250, 240]
    if h > 4:
        img[h - 2, :] = 0
    return img


IMAGES = {
    "pattern": reference_pattern,
    "speckle_97x13": lambda: speckle(97, 13, 4),
    "speckle_1x7": lambda: speckle(1, 7, 5),
    "speckle_5x3": lambda: speckle(5, 3, 6),
    "speckle_256x24": lambda: speckle(256, 24, 7),
}
