"""Layer order of the warm-start U-Net as the library stores its parameters."""


def layer_specs(cin=16, base=16, levels=4):
    """(name, k, cin, cout, relu) of every convolution: encoder levels, decoder
    levels coarse to fine, then the 1^3 output convolution."""
    ch = [base * (2 ** l) for l in range(levels)]
    specs, c = [], cin
    for l in range(levels):
        specs += [(f"enc{l}a", 3, c, ch[l], True), (f"enc{l}b", 3, ch[l], ch[l], True)]
        c = ch[l]
    for l in range(levels - 2, -1, -1):
        specs += [(f"dec{l}a", 3, c + ch[l], ch[l], True), (f"dec{l}b", 3, ch[l], ch[l], True)]
        c = ch[l]
    specs.append(("out", 1, c, cin, False))
    return specs
