"""Transcribes the reference's surface-transport fixtures
(crates/gravlume-reference/fixtures/v3/kerr-blackbody-*.toml) into
surface_transport_fixtures.json: the numbers as strings, unchanged. Run in a
checkout that has the reference next to it; the tests read only the JSON."""
import json
import os
import sys

import tomli

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/crates/gravlume-reference/fixtures/v3"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surface_transport_fixtures.json")

out = {"_source": "Values of the reference's fixtures crates/gravlume-reference/fixtures/v3/"
                  "kerr-blackbody-*.toml (80-digit invariant-transfer and Planck integrals on the "
                  "strict-policy surface hit of fixtures/v2), transcribed as data."}
for name in sorted(os.listdir(SRC)):
    if not name.endswith(".toml"):
        continue
    with open(os.path.join(SRC, name), "rb") as f:
        d = tomli.load(f)
    out[name[:-5]] = {k: d[k] for k in ("base_observation_id", "sample", "surface", "emission",
                                        "transport", "expected", "tolerance")}
with open(OUT, "w") as f:
    json.dump(out, f, indent=1)
    f.write("\n")
print("wrote", OUT, len(out) - 1, "fixtures")
