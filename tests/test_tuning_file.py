"""The committed kernel-variant cache (inverse-solver-gpu_amd/tuning/mi355x.txt)
holds only choices the engine accepts (CPU): one line per (tree depth,
precision, batch class, stage, level), values in the ranges isg_tuning_set
admits -- the split-FP16 near-field kernels (128, 256) and P2M (128) for the
FP32 operator only -- and every (depth, precision, class) it covers names a
near-field and a P2M choice.  A hand edit that slipped a value the engine
rejects would make every context fail to load the cache."""
import collections
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNING = os.path.join(ROOT, "inverse-solver-gpu_amd", "tuning", "mi355x.txt")

STAGES = {"near", "p2m", "m2l", "m2m", "l2l", "leaf"}
M2L_VARIANTS = 10
INTERP_KINDS = 5


def entries():
    out = []
    with open(TUNING) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            depth, prec, cls, stage, level, value = line.split()
            out.append((int(depth), int(prec), int(cls), stage, int(level), int(value)))
    return out


def test_tuning_file_values_admissible():
    seen = set()
    for depth, prec, cls, stage, level, value in entries():
        key = (depth, prec, cls, stage, level)
        assert key not in seen, key
        seen.add(key)
        assert prec in (0, 1) and cls in (0, 1) and stage in STAGES, key
        if stage == "near":
            ok = value in (-1, 16, 32, 64) or (prec == 1 and value in (128, 256))
        elif stage == "p2m":
            ok = value in (-1, 16, 32, 64) or (prec == 1 and value == 128)
        elif stage == "leaf":
            ok = value in (-1, 0, 1)
        elif stage == "m2l":
            ok = -1 <= value < M2L_VARIANTS and 0 <= level < depth
        else:
            ok = -1 <= value < INTERP_KINDS and 0 <= level < depth
        assert ok, (key, value)
        if value == 64 and stage in ("near", "p2m"):
            assert prec == 1, key  # 64-box tiles exist for the FP32 operator only


def test_tuning_file_covers_near_and_p2m():
    by = collections.defaultdict(set)
    for depth, prec, cls, stage, level, value in entries():
        by[(depth, prec, cls)].add(stage)
    assert by, "empty tuning file"
    for key, stages in by.items():
        assert {"near", "p2m"} <= stages, key
