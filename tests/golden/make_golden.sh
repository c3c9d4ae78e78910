#!/usr/bin/env bash
# Regenerates the golden fixtures in this folder from the reference
# pathtracer's own code (run in a container that has the reference checkout).
#   reference_golden.json  outputs of the reference's Embree-free functions,
#                          written by oracle/_ref/golden_harness
#   ior_tables.json        the Au/Al/Fe IOR tables of the reference's
#                          materials/data/*_ior.gltf (data, 300 K)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
REF="${REF:-/root/reference}"
REF="$REF" "$ROOT/oracle/build_ref.sh"
TMP="$(mktemp -d)"
"$ROOT/oracle/_ref/golden_harness" "$REF" "$TMP" > "$HERE/reference_golden.json"
rm -rf "$TMP"
python3 - "$REF" "$HERE/ior_tables.json" <<'EOF'
import json, sys
ref, out = sys.argv[1], sys.argv[2]
tables = {}
for m in ("gold", "aluminium", "iron"):
    d = json.load(open(f"{ref}/materials/data/{m}_ior.gltf"))
    tables[m] = [{"temperature": e["temperature"], "n": e["n"], "k": e["k"]}
                 for e in d["extensions"]["OpenMaterial_ior_data"]["data"]]
json.dump(tables, open(out, "w"))
EOF
python3 -c "import json,sys; d=json.load(open('$HERE/reference_golden.json')); print({k: len(v) for k, v in d.items()})"
