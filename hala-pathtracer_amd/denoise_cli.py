"""The reference's offline denoise step (denoise.py:1-44) with the learned
denoiser in place of `oidnDenoise`: every `<dir>/**/*_color.pfm` that has its
`_albedo.pfm` and `_normal.pfm` beside it gets a `_final.pfm`, denoised on
the GPU by the committed network (denoiser.py).

    python hala-pathtracer_amd/denoise_cli.py            # ./out/, like denoise.py
    python hala-pathtracer_amd/denoise_cli.py --dir out --spp 64

The network takes the sample count as an input; PFM files do not carry it,
so it comes from `--spp`, else from a `<stem>_spp.txt` the headless app
writes beside the images, else 16.
"""
from __future__ import annotations

import argparse
import os
import re
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import denoiser  # noqa: E402
import hpt  # noqa: E402


def find_color_files(directory: str) -> list:
    """denoise.py:7-13: every *_color.pfm under `directory`."""
    out = []
    for root, _dirs, files in os.walk(directory):
        for f in sorted(files):
            if re.search(r"_color.pfm$", f):
                out.append(os.path.join(root, f))
    return sorted(out)


def find_albedo_normal_files(color_file: str):
    """denoise.py:16-23: the matching _albedo.pfm and _normal.pfm, or (None, None)."""
    albedo = color_file.replace("_color.pfm", "_albedo.pfm")
    normal = color_file.replace("_color.pfm", "_normal.pfm")
    if not os.path.exists(albedo) or not os.path.exists(normal):
        return None, None
    return albedo, normal


def samples_of(color_file: str, spp: int | None) -> int:
    if spp:
        return spp
    side = color_file.replace("_color.pfm", "_spp.txt")
    if os.path.exists(side):
        with open(side) as f:
            return max(1, int(f.read().strip() or 16))
    return 16


def _float4(rgb: np.ndarray) -> np.ndarray:
    h, w = rgb.shape[:2]
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = rgb
    return out


def denoise(net: denoiser.UNet, color_file: str, albedo_file: str, normal_file: str, spp: int) -> str:
    """denoise.py:26-30 on the GPU: writes and returns `<stem>_final.pfm`."""
    final_file = color_file.replace("_color.pfm", "_final.pfm")
    imgs = [denoiser.DeviceArray.from_numpy(_float4(hpt.read_pfm(p))) for p in (color_file, albedo_file, normal_file)]
    h, w = imgs[0].shape[:2]
    if imgs[1].shape[:2] != (h, w) or imgs[2].shape[:2] != (h, w):
        raise ValueError(f"{color_file}: the albedo and normal images differ in size")
    out = net.denoise(tuple(imgs), w, h, spp).numpy()
    hpt.write_pfm(final_file, out[..., :3])
    return final_file


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Denoise saved renders with the learned denoiser (replaces oidnDenoise).")
    ap.add_argument("--dir", default="./out/")
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel of the renders (0: from *_spp.txt, else 16)")
    ap.add_argument("--weights", default=denoiser.weights_path())
    a = ap.parse_args(argv)
    net = denoiser.UNet.load(a.weights)
    for color_file in find_color_files(a.dir):
        albedo_file, normal_file = find_albedo_normal_files(color_file)
        if albedo_file is None or normal_file is None:
            continue
        print(f"Color file: {color_file}")
        print(f"Albedo file: {albedo_file}")
        print(f"Normal file: {normal_file}")
        final_file = denoise(net, color_file, albedo_file, normal_file, samples_of(color_file, a.spp))
        print(f"Final file: {final_file}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
