"""Launch-shape autotuner.

Picks, per workload key, the fastest of a list of candidate launch
configurations by timing each on the device, and caches the choice in a JSON
file next to the built library (one entry per key: the candidate and the
measured milliseconds of every candidate).  Keys describe the workload
(kernel family, dtype, size class, robot count); candidates are small dicts
of launch parameters (block size, LDS placement policy, kernel variant, MPM
tile capacity ...).
"""

import hashlib
import json
import os
from pathlib import Path

_DEFAULT = Path(__file__).resolve().parent / "_lib" / "autotune.json"
_LIB = Path(__file__).resolve().parent / "_lib" / "libdiffloco.so"


def library_tag(path=_LIB):
    """Short content hash of the built library: part of every tuning key, so
    a rebuilt library never reuses choices timed on another build."""
    try:
        return hashlib.sha1(Path(path).read_bytes()).hexdigest()[:12]
    except OSError:
        return "none"


class Autotuner:
    def __init__(self, path=None, enabled=True):
        self.path = Path(path) if path is not None else _DEFAULT
        self.enabled = enabled
        self.cache = {}
        if self.path.exists():
            try:
                self.cache = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.cache = {}

    @staticmethod
    def key(**parts):
        return "|".join(f"{k}={parts[k]}" for k in sorted(parts))

    def pick(self, key, candidates, measure, repeat=2, verify=True):
        """Return the candidate with the smallest ``measure(candidate)``
        (milliseconds, best of ``repeat``); cached per key.  Candidates whose
        measurement raises are skipped; with tuning disabled (or one
        candidate) the first candidate is returned untimed.  With ``verify`` a
        cached choice is run once first; if that raises (a stale entry: the
        workload or the library changed under the same key) the entry is
        dropped and the key re-tuned."""
        if key in self.cache:
            want = self.cache[key]["choice"]
            for c in candidates:
                if c == want:
                    if not verify:
                        return c
                    try:
                        measure(c)
                        return c
                    except Exception:
                        del self.cache[key]
                        break
        if not self.enabled or len(candidates) == 1:
            return candidates[0]
        timings = []
        for c in candidates:
            try:
                t = min(measure(c) for _ in range(repeat))
            except Exception as e:          # an invalid shape for this workload
                timings.append({"candidate": c, "error": str(e)[:200]})
                continue
            timings.append({"candidate": c, "ms": t})
        ok = [t for t in timings if "ms" in t]
        if not ok:
            raise RuntimeError(f"autotune {key}: no candidate ran: {timings}")
        best = min(ok, key=lambda t: t["ms"])["candidate"]
        self.cache[key] = {"choice": best, "timings": timings}
        self.save()
        return best

    def save(self):
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(".tmp")
            tmp.write_text(json.dumps(self.cache, indent=1, sort_keys=True))
            os.replace(tmp, self.path)
        except OSError:
            pass
