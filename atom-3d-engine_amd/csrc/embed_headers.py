"""Build step: embed the device headers into the library as C++ raw strings, so the
run-time compiler (hipRTC) can include them without the source tree being present."""
import os
import sys

out, names = sys.argv[1], sys.argv[2:]
delim = "MM3EHDR"
lines = ["// generated by embed_headers.py — device headers for the scene code generator",
         '#include "codegen.h"', "namespace mm3e {", "const EmbeddedHeader kEmbeddedHeaders[] = {"]
for n in names:
    text = open(n).read()
    assert (")" + delim + '"') not in text
    lines.append('  {"%s", R"%s(%s)%s"},' % (os.path.basename(n), delim, text, delim))
lines += ["};", "const int kEmbeddedHeaderCount = %d;" % len(names), "}  // namespace mm3e", ""]
open(out, "w").write("\n".join(lines))
