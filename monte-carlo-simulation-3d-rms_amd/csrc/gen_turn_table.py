"""Writes rms_turn_table.h: sin and cos of 2*pi*k/256, k = 0..255, as the
float32 roundings of the C library's double sin/cos (the values the host
restatement computes the same way).  Run from this directory:
python gen_turn_table.py > rms_turn_table.h"""
import math

import numpy as np

print("// rms_turn_table.h -- sin and cos of 2*pi*k/256 (k = 0..255) as float32, the")
print("// roundings of the C library's double sin/cos.  Written by gen_turn_table.py.")
print("#pragma once")
print("")
print("namespace rms {")
print("")
print("// {sin, cos} pairs")
print("__device__ const float2 turn_table[256] = {")
for k in range(256):
    a = 2 * math.pi * k / 256
    s = float(np.float32(math.sin(a)))
    c = float(np.float32(math.cos(a)))
    print(f"    {{{s.hex()}f, {c.hex()}f}},")
print("};")
print("")
print("}  // namespace rms")
