"""Write tests/golden/rex_small.npz: a seeded rex case (ZB, three frames with
their track values and model clocks, a track file, the envelope field) and the
rex file the reference's own writer makes of it (oracle/_ref/rexref64, built
from the reference's rex.c, pack.c and their helpers).  Run from the
repository root after `make -C oracle`."""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import rexutil  # noqa: E402


def main():
    rng = np.random.default_rng(2025)
    im, jm = 33, 27
    _, zb = rexutil.frame_fields(rng, im, jm)
    kinds = ["mixed", "dry", "mixed"]
    hbs = np.stack([rexutil.frame_fields(rng, im, jm, k, zb)[0] for k in kinds])
    track = np.array([[25.12345 + k, 80.98765 - k, 12.345, 271.115, 55.555, 17.125] for k in range(3)])
    clock = np.array([1756685400.0, 1756686000.0, 1756686600.25])
    hb_env = rexutil.frame_fields(rng, im, jm, "mixed", zb)[0]
    root = tempfile.mkdtemp()
    trk = os.path.join(root, "t.trk")
    rexutil.trk_2013(trk, 48, rng, ocean="0.3", lake="-1.2", nrows=60)
    spec = os.path.join(root, "spec.bin")
    comment, abbrev = "FileVersion=2013", "tny"
    rexutil.write_spec(spec, zb, [(track[k], clock[k], hbs[k]) for k in range(3)], trk, hb_env, comment, abbrev)
    ref = rexutil.reference_rex(spec, os.path.join(root, "ref.rex"))
    np.savez_compressed(os.path.join(HERE, "rex_small.npz"), zb=zb, hb=hbs, track=track, clock=clock,
                        hb_env=hb_env, trk=np.frombuffer(open(trk, "rb").read(), np.uint8),
                        ref_rex=np.frombuffer(ref, np.uint8), comment=comment, abbrev=abbrev)
    print("rex_small.npz: %d byte rex file, %d frames" % (len(ref), len(hbs)))


if __name__ == "__main__":
    main()
