"""Embeds text files as one C++ raw string constant (build step).

    python3 embed.py OUT NAME IN...

The inputs are concatenated in order; an `#include "X"` line naming another
input is dropped, since X's text is already in the string."""
import os
import sys

dst, name, srcs = sys.argv[1], sys.argv[2], sys.argv[3:]
names = {os.path.basename(s) for s in srcs}
text = ""
for src in srcs:
    for line in open(src).read().splitlines(keepends=True):
        s = line.strip()
        if s.startswith('#include "') and s[len('#include "'):-1] in names:
            continue
        text += line
delim = "SIMKEMBED"
assert f"){delim}\"" not in text
with open(dst, "w") as f:
    f.write(f"// generated from {', '.join(srcs)} by embed.py\n")
    f.write(f"static const char {name}[] = R\"{delim}(\n{text}){delim}\";\n")
