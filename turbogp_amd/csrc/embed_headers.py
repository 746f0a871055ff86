"""Writes the C++ include file that embeds the device headers for run-time
compilation (jit.hip): one raw string per header, under the name its
#include uses in the embedded set (the public header's relative path is
rewritten to its bare name, since hipRTC resolves the names given to it)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADERS = [("tgp_internal.h", "tgp_internal.h"), ("tile_engine.h", "tile_engine.h"),
           ("plan.h", "plan.h"), ("spec.h", "spec.h"), ("wave_agg.h", "wave_agg.h"),
           ("spec_kernels.h", "spec_kernels.h"),
           ("turbogp_gfx950.h", os.path.join("..", "..", "include", "turbogp_gfx950.h"))]
DELIM = "tgpjit"


def main(out):
    parts = []
    for name, rel in HEADERS:
        with open(os.path.join(HERE, rel)) as f:
            text = f.read().replace('#include "../../include/turbogp_gfx950.h"',
                                    '#include "turbogp_gfx950.h"')
        assert ")" + DELIM + '"' not in text, name
        parts.append((name, text))
    with open(out + ".tmp", "w") as f:
        f.write("// Generated by embed_headers.py from the headers it names; do not edit.\n")
        f.write(f"static const int kJitNumHeaders = {len(parts)};\n")
        f.write("static const char* const kJitHeaderNames[] = {")
        f.write(", ".join(f'"{n}"' for n, _ in parts) + "};\n")
        f.write("static const char* const kJitHeaderText[] = {\n")
        for _, text in parts:
            f.write(f'R"{DELIM}({text}){DELIM}",\n')
        f.write("};\n")
    os.replace(out + ".tmp", out)


if __name__ == "__main__":
    main(sys.argv[1])
