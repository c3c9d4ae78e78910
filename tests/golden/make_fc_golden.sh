#!/usr/bin/env bash
# Regenerates tests/golden/fc_golden.json: what the reference pathtracer's own
# false-colour subjects (validation/FCSubject.h, FC_VALIDATION build) paint
# for scripted arguments, written by oracle/_ref/fc_harness (built from the
# reference sources where they lie by oracle/build_ref.sh).  Run in a
# container that has the reference checkout.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
REF="${REF:-/root/reference}"
REF="$REF" "$ROOT/oracle/build_ref.sh"
"$ROOT/oracle/_ref/fc_harness" > "$HERE/fc_golden.json"
python3 -c "import json; d=json.load(open('$HERE/fc_golden.json')); print(sorted(d))"
