"""Writes the golden fixtures under tests/golden/.

measurement_known_answers.json: the post-measurement amplitudes that
MeasurementsTests.cpp:75-85 and :138-149 assert for the state
1/2|00> - i/2|01> + 1/sqrt(2)|11> (MeasurementsTests.cpp:34-43), with the
uniform u that selects each branch under prob = 1 - u and the first index
with prob <= cdf (cdf = 0.25, 0.5, 0.5, 1.0 over states 0..3).
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    r3, r2 = math.sqrt(3.0), math.sqrt(2.0)
    cases = [
        # qubit 0 measured 1: (0, -i/sqrt3, 0, sqrt2/sqrt3)      (MeasurementsTests.cpp:75-76)
        {"qubit": 0, "u": 0.1, "outcome": 1, "amplitudes": [[0, 0], [0, -1 / r3], [0, 0], [r2 / r3, 0]]},
        {"qubit": 0, "u": 0.6, "outcome": 1, "amplitudes": [[0, 0], [0, -1 / r3], [0, 0], [r2 / r3, 0]]},
        # qubit 0 measured 0: (1, 0, 0, 0)                          (MeasurementsTests.cpp:84-85)
        {"qubit": 0, "u": 0.9, "outcome": 0, "amplitudes": [[1, 0], [0, 0], [0, 0], [0, 0]]},
        # qubit 1 measured 1: (0, 0, 0, 1)                          (MeasurementsTests.cpp:138-139)
        {"qubit": 1, "u": 0.1, "outcome": 1, "amplitudes": [[0, 0], [0, 0], [0, 0], [1, 0]]},
        # qubit 1 measured 0: (sqrt2/2, -i sqrt2/2, 0, 0)           (MeasurementsTests.cpp:147-148)
        {"qubit": 1, "u": 0.6, "outcome": 0, "amplitudes": [[0.5 * r2, 0], [0, -0.5 * r2], [0, 0], [0, 0]]},
    ]
    with open(os.path.join(HERE, "measurement_known_answers.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    main()
