"""Writes known_answers.json: closed-form known-answer vectors of the reconstruct path.

Values are the published closed forms restated in the survey of the reference (FNV fold60 keys,
table sizing of hashmap_cuda.cu:341-360, exact B-spline integral tables) — not outputs of any
implementation in this repository. F, G, P, Pinv/Qinv are the survey's listed values; Q is
m * (B((ra+1)/m) - B(ra/m)) (the survey's formula int_{ra}^{ra+1} B'(x/m) dx; its listed Q_2,
Q_4 carry an extra factor m and are not used).
"""
import json
import os

KA = {
    "fnv_fold60": [
        {"coords": [0, 0, 0], "key": "0x094d12186c0f2fba"},
        {"coords": [1, 2, 3], "key": "0x00aa6218672cf5a6"},
        {"coords": [-1, 0, 5], "key": "0x0bdb7f2a8f08da19"},
        {"coords": [127, 127, 127], "key": "0x032734196f09ea9e"},
        {"coords": [0, 1], "key": "0x08328707b4eb6e3a"},
    ],
    "sizing": [
        {"n": 100, "table_size": 512, "num_buckets": 1, "evict_bound": 56},
        {"n": 2047, "table_size": 8192, "num_buckets": 16, "evict_bound": 88},
        {"n": 2048, "table_size": 4096, "num_buckets": 8, "evict_bound": 88},
        {"n": 86415, "table_size": 262144, "num_buckets": 512, "evict_bound": 136},
        {"n": 3115282, "table_size": 8388608, "num_buckets": 16384, "evict_bound": 176},
    ],
    "tables": {
        "F": {"1": ["1/30", "13/15", "11/5", "13/15", "1/30"],
              "2": ["1/120", "29/120", "49/40", "101/40", "101/40", "49/40", "29/120", "1/120"],
              "4": ["1/480", "29/480", "5/16", "13/16", "249/160", "381/160", "23/8", "23/8", "381/160",
                    "249/160", "13/16", "5/16", "29/480", "1/480"]},
        "G": {"1": ["-2/3", "-4/3", "4", "-4/3", "-2/3"],
              "2": ["-1/3", "-5/3", "-1", "3", "3", "-1", "-5/3", "-1/3"],
              "4": ["-1/6", "-5/6", "-1", "-1", "-1/2", "3/2", "2", "2", "3/2", "-1/2", "-1", "-1", "-5/6",
                    "-1/6"]},
        "P": {"1": ["1/3", "4/3", "1/3"], "2": ["1/12", "7/12", "4/3", "4/3", "7/12", "1/12"],
              "4": ["1/48", "7/48", "19/48", "37/48", "29/24", "35/24", "35/24", "29/24", "37/48", "19/48",
                    "7/48", "1/48"]},
        "Q": {"1": ["-1", "0", "1"], "2": ["-1/2", "-3/2", "-1", "1", "3/2", "1/2"],
              "4": ["-1/4", "-3/4", "-5/4", "-7/4", "-3/2", "-1/2", "1/2", "3/2", "7/4", "5/4", "3/4",
                    "1/4"]},
        "Pinv": {"2": ["1/3", "5/3", "5/3", "1/3"]},
        "Qinv": {"2": ["-1", "-1", "1", "1"]},
    },
    # 1-D quadratic B-spline Galerkin stencils (textbook); the basis here is 2 N2, so x4.
    "stencil_1d": {"stiffness": ["-1/6", "-1/3", "1", "-1/3", "-1/6"],
                   "mass": ["1/120", "13/60", "11/20", "13/60", "1/120"]},
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump(KA, f, indent=1)
    print(out)
