#!/bin/bash
# Rebuilds tests/golden/trace_gaussians.txt from gen_trace_gaussians.cpp.
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
g++ -O2 -std=c++17 -o "$tmp/gen" "$here/gen_trace_gaussians.cpp"
"$tmp/gen" 4000 > "$here/trace_gaussians.txt"
rm -rf "$tmp"
