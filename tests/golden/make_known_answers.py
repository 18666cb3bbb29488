"""Write reference_known_answers.json: the known answers that the
reference's own tests and micro-benchmarks assert for the ray-casting path,
as data.  Each entry cites the reference file:line it comes from.  Run:

    python tests/golden/make_known_answers.py
"""
import json
import math
import os

sc = math.sqrt(2.0) / 2.0

data = {
    "source": "shamazmazum/voxvision tests/tests.c and benchmarks/voxtrees",
    # tests.c:957-959
    "tolerances": {"precise_check": 1e-5, "approx_check": 0.5},
    # tests.c:303-363 (quat_mul)
    "quat_mul": [
        {"q1": [0, 1, 0, 0], "q2": [0, 0, 1, 0], "expect": [0, 0, 0, 1]},
        {"q1": [0, 0, 1, 0], "q2": [0, 0, 0, 1], "expect": [0, 1, 0, 0]},
        {"q1": [0, 0, 0, 1], "q2": [0, 1, 0, 0], "expect": [0, 0, 1, 0]},
        {"q1": [0, 0, 1, 0], "q2": [0, 1, 0, 0], "expect": [0, 0, 0, -1]},
        {"q1": [0, 0, 0, 1], "q2": [0, 0, 1, 0], "expect": [0, -1, 0, 0]},
        {"q1": [0, 1, 0, 0], "q2": [0, 0, 0, 1], "expect": [0, 0, -1, 0]},
        {"q1": [4, 3, 2, 1], "q2": [1, 0, 0, 0], "expect": [4, 3, 2, 1]},
        {"q1": [1, 0, 0, 0], "q2": [4, 3, 2, 1], "expect": [4, 3, 2, 1]},
        {"q1": [4, 1, 2, 3], "q2": [2, 0, 2, 0], "expect": [4, -4, 12, 8]},
    ],
    # tests.c:85-104 (identity), :106-121 (rotation around itself),
    # :123-146 (change axis); expect "v" = the input vector
    "rotate_vector": [
        {"q": [1, 0, 0, 0], "v": [1, 2, 4], "expect": "v"},
        {"q": [1, 0, 0, 0], "v": [-321, 22, 124], "expect": "v"},
        {"q": [1, 0, 0, 0], "v": [-321, 22, -124], "expect": "v"},
        {"q": [math.cos(0.1), math.sin(0.1), 0, 0], "v": [1, 0, 0],
         "expect": "v"},
        {"q": [math.cos(0.1), 0, math.sin(0.1), 0], "v": [0, 1, 0],
         "expect": "v"},
        {"q": [math.cos(0.1), 0, 0, math.sin(0.1)], "v": [0, 0, 1],
         "expect": "v"},
        {"q": [sc, 0, sc, 0], "v": [1, 0, 0], "expect": [0, 0, -1]},
        {"q": [sc, 0, 0, sc], "v": [0, 1, 0], "expect": [-1, 0, 0]},
        {"q": [sc, sc, 0, 0], "v": [0, 0, 1], "expect": [0, -1, 0]},
    ],
    # tests.c:661-704 (camera generic test), :706-724 (look_at)
    "camera": {"window": [100, 100], "pixel": [50, 50],
               "center_ray": [0, 1, 0], "after_z_rotation": [-1, 0, 0],
               "look_at_position": [-100, 20, -20]},
    # tests.c:214-237: ball of radius 50 centred at the origin
    "ball50": {"voxels": 523155},
    # tests.c:600-659 (search regression of commit 676d50c)
    "search_676d50c": {
        "voxels": [[0, 0, 0], [0, 0, -1], [0, 0, -2], [0, 0, 1], [0, 0, 2],
                   [-1, 0, 0], [-2, 0, 0], [-3, 0, 0], [-4, 0, 0],
                   [1, 0, 0], [2, 0, 0], [3, 0, 0], [4, 0, 0],
                   [0, -1, 0], [0, -2, 0], [0, -3, 0], [0, -4, 0],
                   [0, 1, 0], [0, 2, 0], [0, 3, 0], [0, 4, 0]],
        "origin": [-2, -2, 0], "dir": [1, 1, -1], "expect_hit": True},
    # benchmarks/voxtrees/hit-box.c:5-31
    "hit_box_benchmark": {"voxel": [0, 10, 0], "hits_per_10": 1,
                          "hit_point": [0, 10, 0]},
    # benchmarks/voxtrees/ball-intersection.c:5-28
    "ball_benchmark": {"voxel": [0, 0, 0], "radius": 1000.0,
                       "collisions_per_200": 200},
}

out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "reference_known_answers.json")
with open(out, "w") as f:
    json.dump(data, f, indent=1)
print(out)
