#!/bin/sh
# Regenerates the XML goldens from the reference's own loader/writer
# (src/gpu/shared/{input,output,common}.cpp + tinyxml2), built from the
# reference sources by `make -C oracle ref` into oracle/_ref/. The parse is
# stored as its scalar fields plus length, sum and SHA-256 of the K and DE
# lists (values formatted with %.17g).
set -e
d=$(dirname "$0")
make -C "$d/../../oracle" ref >/dev/null
"$d/../../oracle/_ref/ref_xml" load "$d/test_files/input.xml" "$d/test_files/data_20171018_175nm_500nm.xml" \
    > /tmp/ref_xml_parse_full.json
python3 - "$d/ref_xml_parse.json" <<'PY'
import hashlib, json, sys
full = json.load(open("/tmp/ref_xml_parse_full.json"))
out = {k: v for k, v in full.items() if k not in ("K", "DE", "number_of_frames")}
for k in ("K", "DE", "number_of_frames"):
    vals = full[k]
    txt = ",".join("%.17g" % v for v in vals)
    out[k] = {"n": len(vals), "sum": float(sum(vals)), "sha256": hashlib.sha256(txt.encode()).hexdigest(),
              "min": min(vals), "max": max(vals)}
json.dump(out, open(sys.argv[1], "w"), indent=1, sort_keys=True)
PY
"$d/../../oracle/_ref/ref_xml" write "$d/ref_output.xml"
