"""Embeds a text file as a C++ raw string constant (build step)."""
import sys

src, dst, name = sys.argv[1:4]
text = open(src).read()
delim = "SIMKEMBED"
assert f"){delim}\"" not in text
with open(dst, "w") as f:
    f.write(f"// generated from {src} by embed.py\n")
    f.write(f"static const char {name}[] = R\"{delim}(\n{text}){delim}\";\n")
