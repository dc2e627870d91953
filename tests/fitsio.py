"""Minimal reader for the primary HDU of the simple FITS images used as golden fixtures."""
import numpy as np


def read_fits(path):
    with open(path, 'rb') as f:
        data = f.read()
    hdr = {}
    pos = 0
    while True:
        card = data[pos:pos + 80].decode('ascii')
        pos += 80
        key = card[:8].strip()
        if key == 'END':
            break
        if card[8:10] == '= ':
            hdr[key] = card[10:].split('/')[0].strip()
    pos = ((pos + 2879) // 2880) * 2880
    bitpix = int(hdr['BITPIX'])
    n1, n2 = int(hdr['NAXIS1']), int(hdr['NAXIS2'])
    dt = {-32: '>f4', -64: '>f8', 32: '>i4', 16: '>i2'}[bitpix]
    arr = np.frombuffer(data[pos:pos + n1 * n2 * abs(bitpix) // 8], dtype=dt).reshape(n2, n1)
    return arr.astype(np.float64)
