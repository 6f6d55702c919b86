"""The SiSter dataset on the MI355X path (python/extract_dataset.py:9-183,
README.md:23-61).

Two layouts are read:

  * the raw capture folders, as the reference's SisterDataset reads them
    (extract_dataset.py:9-96): a scenes folder of <object>/<prefix>_<level>_
    <baseline>/00000_<view>.png (levels "0", "1", "2" are camera distances of
    1, 5 and 10 cm, LEVEL_TO_HEIGHT_MAP at :10) plus predictions under
    output/<name>.png, and a ground-truth folder of <object>_<level>.exr files;
  * the exported layout the README documents (README.md:40-61) and export()
    writes (extract_dataset.py:148-183): <object>/<distance>/<baseline>mm/
    <view>.png and <object>/<distance>/gt_depth.exr.

Both give a rig's five views in the reference's constructor order (center,
right, top, left, bottom) as the class takes them (BGR, converted to gray
like cvtColor(BGR2GRAY)), and the depth ground truth in metres (float32).
run() computes a rig's three maps through the HIP library and writes them as
16-bit PNGs beside the rig (the reference's compute_disp outputs), or under
the raw layout's output/ folder, where load_prediction reads them.

    python -m sister_amd.dataset export --scenes_folder S --gt_folder G --output_folder O
    python -m sister_amd.dataset run --root O --disp 64 [--object arduino] [--modes 7]
"""
from __future__ import annotations

import argparse
import os
import shutil
from pathlib import Path

import numpy as np

from . import imageio

VIEWS = ("center", "right", "top", "left", "bottom")   # constructor order (hpp:22)


class SisterDataset:
    """The raw capture folders (extract_dataset.py:9-96): same method names
    and arguments."""

    LEVEL_TO_HEIGHT_MAP = {"0": "1cm", "1": "5cm", "2": "10cm"}

    def __init__(self, scenes_folder, gt_folder) -> None:
        self.scenes_folder = Path(scenes_folder)
        self.gt_folder = Path(gt_folder)
        self.scenes_map = {}
        gts_map = self.build_gts_map(self.gt_folder)
        self.object_names = set()
        for f in self.subfolders(self.scenes_folder):
            self.scenes_map[f.name] = {"images": self.build_levels_map(f), "gt": gts_map.get(f.name)}
            if f.name in gts_map:
                self.object_names.add(f.name)

    def images_folder(self, object_name, level="1", baseline="100"):
        return self.scenes_map[object_name]["images"][level][baseline]

    def object_baselines(self, object_name, level):
        return list(self.scenes_map[object_name]["images"][level].keys())

    def load_image(self, object_name, level="1", baseline="100", direction="center"):
        return imageio.read_png(Path(self.images_folder(object_name, level, baseline)) / f"00000_{direction}.png",
                                anydepth=False)

    def load_views(self, object_name, level="1", baseline="100"):
        return [self.load_image(object_name, level, baseline, v) for v in VIEWS]

    def load_prediction(self, object_name, level="1", baseline="100", prediction_name="FULL_mccnn_refine"):
        return imageio.read_png(Path(self.images_folder(object_name, level, baseline)) / "output" /
                                f"{prediction_name}.png")

    def gt_levels(self, object_name):
        return list(self.scenes_map[object_name]["gt"].keys())

    def gt_file(self, object_name, level="1"):
        return self.scenes_map[object_name]["gt"][level]

    def load_gt(self, object_name, level="1"):
        return imageio.read_exr(self.gt_file(object_name, level))

    @staticmethod
    def subfolders(folder):
        return sorted(x for x in Path(folder).glob("*") if x.is_dir())

    @staticmethod
    def innerfiles(folder):
        return sorted(x for x in Path(folder).glob("*") if not x.is_dir())

    def build_levels_map(self, folder):
        # <prefix>_<level>_<baseline> (extract_dataset.py:73-83)
        levels_map = {}
        for sub in self.subfolders(folder):
            _, level_id, baseline = sub.stem.split("_")
            levels_map.setdefault(level_id, {})[baseline] = sub
        return levels_map

    def build_gts_map(self, folder):
        # <object name with underscores>_<level>.exr (extract_dataset.py:85-96)
        gts_map = {}
        for f in self.innerfiles(folder):
            chunks = f.stem.split("_")
            gts_map.setdefault("_".join(chunks[:-1]), {})[chunks[-1]] = f
        return gts_map


def export(scenes_folder, gt_folder, output_folder):
    """Re-layout of the raw folders into the documented one
    (extract_dataset.py:148-183): views copied without their 00000_ prefix
    into <object>/<distance>/<baseline>mm/, the level's ground truth into
    <object>/<distance>/gt_depth.exr."""
    ds = SisterDataset(scenes_folder=scenes_folder, gt_folder=gt_folder)
    out = Path(output_folder)
    out.mkdir(parents=True, exist_ok=True)
    for obj in sorted(ds.object_names):
        for level in sorted(ds.gt_levels(obj)):
            level_folder = out / obj / SisterDataset.LEVEL_TO_HEIGHT_MAP[level]
            level_folder.mkdir(parents=True, exist_ok=True)
            for baseline in ds.object_baselines(obj, level):
                bf = level_folder / f"{baseline}mm"
                bf.mkdir(parents=True, exist_ok=True)
                for src in sorted(Path(ds.images_folder(obj, level, baseline)).glob("*.png")):
                    shutil.copy(src, bf / src.name.replace("00000_", ""))
            shutil.copy(ds.gt_file(obj, level), level_folder / "gt_depth.exr")
    return out


class ExportedDataset:
    """The documented layout (README.md:40-61):
    <object>/<distance>/<baseline>mm/<view>.png, <object>/<distance>/gt_depth.exr."""

    def __init__(self, root):
        self.root = Path(root)

    def objects(self):
        return sorted(p.name for p in self.root.iterdir() if p.is_dir())

    def distances(self, obj):
        return sorted(p.name for p in (self.root / obj).iterdir() if p.is_dir())

    def baselines(self, obj, distance):
        return sorted(p.name for p in (self.root / obj / distance).iterdir() if p.is_dir() and p.name.endswith("mm"))

    def rigs(self):
        """Every (object, distance, baseline) with all five views present."""
        out = []
        for o in self.objects():
            for d in self.distances(o):
                for b in self.baselines(o, d):
                    f = self.root / o / d / b
                    if all((f / f"{v}.png").exists() for v in VIEWS):
                        out.append((o, d, b))
        return out

    def rig_folder(self, obj, distance, baseline):
        return self.root / obj / distance / baseline

    def load_views(self, obj, distance, baseline):
        f = self.rig_folder(obj, distance, baseline)
        return [imageio.read_png(f / f"{v}.png", anydepth=False) for v in VIEWS]

    def load_gt(self, obj, distance):
        return imageio.read_exr(self.root / obj / distance / "gt_depth.exr")


def baseline_mm(name):
    """'025mm' or '025' -> 25.0"""
    return float(name[:-2] if name.endswith("mm") else name)


def depth_to_disparity(depth_m, baseline, focal_px):
    """Disparity in pixels of a depth map in metres for a rig baseline in mm
    and a focal length in pixels: d = f B / Z (0 where the depth is not
    positive or not finite).  The dataset does not ship the intrinsics; the
    focal length is the caller's."""
    z = np.asarray(depth_m, np.float64)
    ok = np.isfinite(z) & (z > 0)
    d = np.zeros(z.shape, np.float64)
    d[ok] = focal_px * (baseline_mm(baseline) if isinstance(baseline, str) else baseline) * 1e-3 / z[ok]
    return d.astype(np.float32)


def _to_gray(v):
    from . import bgr_to_gray
    return bgr_to_gray(v)


def compute(views, disp_count, modes=7, context=None):
    """compute_disparities on the HIP path for one rig of dataset views
    (BGR or gray): {"multiview", "horizontal", "vertical"} uint16 maps
    (255 d), plus "disp" and "subpix" when the multiview mode is asked."""
    from . import Context
    ctx = context or Context()
    gray = [_to_gray(v) for v in views]
    try:
        return ctx.compute_rig(gray, int(disp_count), modes=modes)
    finally:
        if context is None:
            ctx.close()


def run(root, disp_count, objects=None, modes=7, out_name="sister", context=None, log=print):
    """Every rig of an exported dataset through the HIP path: writes
    <rig folder>/disp_{multiview,horizontal,vertical}.png (16-bit, 255 d, as
    compute_disp writes them) and returns [(object, distance, baseline,
    seconds)]."""
    import time
    from . import Context
    ds = ExportedDataset(root)
    ctx = context or Context()
    done = []
    try:
        for o, d, b in ds.rigs():
            if objects and o not in objects:
                continue
            t0 = time.perf_counter()
            r = compute(ds.load_views(o, d, b), disp_count, modes, ctx)
            dt = time.perf_counter() - t0
            f = ds.rig_folder(o, d, b)
            for k in ("multiview", "horizontal", "vertical"):
                if k in r:
                    imageio.write_png(f / f"disp_{k}.png", r[k])
            done.append((o, d, b, dt))
            log(f"{o}/{d}/{b}: {dt * 1e3:.1f} ms")
    finally:
        if context is None:
            ctx.close()
    return done


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m sister_amd.dataset")
    sub = ap.add_subparsers(dest="cmd", required=True)
    e = sub.add_parser("export", help="raw capture folders -> documented layout")
    e.add_argument("--scenes_folder", default="objects_full_scenes")
    e.add_argument("--gt_folder", default="objects_full_scenes_GT/objects_full_scenes_gtmanual")
    e.add_argument("--output_folder", required=True)
    r = sub.add_parser("run", help="disparity maps of every rig of an exported dataset")
    r.add_argument("--root", required=True)
    r.add_argument("--disp", type=int, default=64)
    r.add_argument("--object", action="append")
    r.add_argument("--modes", type=int, default=7)
    a = ap.parse_args(argv)
    if a.cmd == "export":
        export(a.scenes_folder, a.gt_folder, a.output_folder)
    else:
        run(a.root, a.disp, a.object, a.modes)


if __name__ == "__main__":
    main()
