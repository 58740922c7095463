"""The batched autotuner's header writer (CPU only): the tables it emits are
the ones batched.hip reads, with the chosen values in [float, double]
order."""
import re

from mat_amd import autotune_batched as at


def test_header_tables():
    def side(nbs, per_cu, order):
        return {"bins": [{"bin": list(b), "ms": {8: 1.0, 16: 0.5}, "nb": nb}
                         for b, nb in zip(at.BINS, nbs)],
                "per_cu": per_cu, "order": order, "schedules": {"1/0": 1.0, "2/1": 0.9}}
    text = at.header({"float": side([16, 16, 8, 16], 3, 0), "double": side([8, 16, 16, 16], 1, 2)},
                     1500, 5000, 3)
    assert "#pragma once" in text
    nb = re.search(r"kTunedNb\[2\]\[4\] = \{\s*/\* float  \*/ \{([^}]*)\},\s*/\* double \*/ \{([^}]*)\}",
                   text)
    assert nb.group(1).replace(" ", "") == "16,16,8,16"
    assert nb.group(2).replace(" ", "") == "8,16,16,16"
    assert "kTunedPerCu[2] = {3, 1};" in text
    assert "kTunedOrder[2] = {0, 2};" in text
