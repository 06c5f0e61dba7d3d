// Oracle: colour conversions, compressed colour leaf decoder/encoder, colour tree.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <algorithm>

namespace orc {

// color_utils.h:9-20
u16 float3_to_rgb565(F3 c) {
    const float r = tclamp(c.x, 0.f, 1.f), g = tclamp(c.y, 0.f, 1.f), b = tclamp(c.z, 0.f, 1.f);
    return (u16)(((u16)(r * 31.0f) << 0) | ((u16)(g * 63.0f) << 5) | ((u16)(b * 31.0f) << 11));
}
// color_utils.h:20-26
F3 rgb565_to_float3(u16 c) {
    return {(float)((c >> 0) & 0x1F) / 31.0f, (float)((c >> 5) & 0x3F) / 63.0f, (float)((c >> 11) & 0x1F) / 31.0f};
}
// color_utils.h:28-37
u32 float3_to_rgb888(F3 c) {
    const float r = tclamp(c.x, 0.f, 1.f), g = tclamp(c.y, 0.f, 1.f), b = tclamp(c.z, 0.f, 1.f);
    return ((u32)(r * 255.0f) << 0) | ((u32)(g * 255.0f) << 8) | ((u32)(b * 255.0f) << 16) | 0xff000000u;
}
// color_utils.h:38-44
F3 rgb888_to_float3(u32 c) {
    return {(float)((c >> 0) & 0xFF) / 255.0f, (float)((c >> 8) & 0xFF) / 255.0f, (float)((c >> 16) & 0xFF) / 255.0f};
}
// color_utils.h:46-51
F3 rgb101210_to_float3(u32 c) {
    return {(float)((c >> 0) & 0x3FF) / 1023.0f, (float)((c >> 10) & 0xFFF) / 4095.0f, (float)((c >> 22) & 0x3FF) / 1023.0f};
}
// color_utils.h:52-60
u32 float3_to_rgb101210(F3 c) {
    const float r = tclamp(c.x, 0.f, 1.f), g = tclamp(c.y, 0.f, 1.f), b = tclamp(c.z, 0.f, 1.f);
    return ((u32)(r * 1023.0f) << 0) | ((u32)(g * 4095.0f) << 10) | ((u32)(b * 1023.0f) << 22);
}

// variable_weight_size_colors.h:114-121, color_utils.h:74-77
float CompressedColor::get_weight() const {
    if (bitsPerWeight == 0) return 0.f;
    return (float)weight / (float)((1 << bitsPerWeight) - 1);
}
// variable_weight_size_colors.h:144-154; lerp(a,b,f) = a*(1-f) + b*f (cuda_math.h:22)
F3 CompressedColor::get_color() const {
    if (bitsPerWeight == 0) return rgb101210_to_float3(colorBits);
    const F3 a = min_color(), b = max_color();
    const float f = get_weight();
    return a * (1 - f) + b * f;
}

// Block header fields (variable_weight_size_colors.h:10-53)
static inline u32 hdr_color_index(u32 h) { return h & (kColorsPerMacro - 1); }
static inline u32 hdr_weight_offset(u32 h) { return h >> 16; }
static inline u32 hdr_bpw(u32 h) { return hdr_weight_offset(h) == 0xFFFF ? 0u : ((h >> 14) & 3u) + 1u; }
static inline u32 make_block_header(u32 weightOffset, u32 bpw, u32 index) {
    if (bpw == 0) weightOffset = 0xFFFF;
    else bpw--;
    return (weightOffset << 16) | (bpw << 14) | index;
}

// color_utils.h:112-148 (byte-swapped storage branch): the weight bit stream is MSB-first;
// read a big-endian 16-bit window starting at byte bitPtr/8.
u8 extract_bits(u32 bits, const std::vector<u32>& words, u64 bitPtr) {
    if (bits == 0) return 0;
    const u64 byte = bitPtr / 8;
    auto byte_at = [&](u64 b) -> u32 {
        if (b / 4 >= words.size()) return 0u;  // only reached for bits that are masked out
        return (words[b / 4] >> (8 * (b % 4))) & 0xFFu;
    };
    const u32 dst = (byte_at(byte) << 8) | byte_at(byte + 1);
    const u32 shift = 16 - bits - (u32)(bitPtr % 8);
    return (u8)((dst >> shift) & ((1u << bits) - 1));
}

// variable_weight_size_colors.h:255-262
bool LeafView::is_valid_index(u64 idx) const {
    if (is_shared()) idx += offset;
    return 2 * idx / kColorsPerMacro < d->macro.size();
}
// variable_weight_size_colors.h:331-341
u64 LeafView::macro_last_block(u64 m) const {
    if (2 * (m + 1) < d->macro.size()) return d->macro[2 * (m + 1)] - 1;
    return d->blocks.size() - 1;
}
// variable_weight_size_colors.h:343-371 (search for the block starting at the local index;
// on exit without a match the position lands on the last block starting before it)
u32 LeafView::binary_search_blocks(u64 colorIndex) const {
    const u32 local = (u32)(colorIndex % kColorsPerMacro);
    const u32 m = (u32)(colorIndex / kColorsPerMacro);
    u32 lo = (u32)d->macro[2 * m];
    u32 hi = (u32)macro_last_block(m);
    u32 pos = (lo + hi) / 2;
    u32 h = (u32)d->blocks[pos];
    while (hdr_color_index(h) != local && lo <= hi) {
        if (hdr_color_index(h) > local) hi = pos - 1;
        else lo = pos + 1;
        pos = (lo + hi) / 2;
        h = (u32)d->blocks[pos];
    }
    return pos;
}
// variable_weight_size_colors.h:373-403
CompressedColor LeafView::color_for_block(u32 block, u64 colorIndex) const {
    const u32 local = (u32)(colorIndex % kColorsPerMacro);
    const u32 m = (u32)(colorIndex / kColorsPerMacro);
    const u64 b = d->blocks[block];
    const u32 h = (u32)b;
    CompressedColor c;
    c.bitsPerWeight = (u8)hdr_bpw(h);
    c.colorBits = (u32)(b >> 32);
    if (c.bitsPerWeight != 0) {
        const u64 macroWeight = d->macro[2 * m + 1];
        const u32 blockWeight = hdr_weight_offset(h);
        const u32 localWeight = (local - hdr_color_index(h)) * c.bitsPerWeight;
        c.weight = extract_bits(c.bitsPerWeight, d->weights, macroWeight + blockWeight + localWeight);
    }
    return c;
}
// variable_weight_size_colors.h:405-413
CompressedColor LeafView::get_color(u64 colorIndex) const {
    if (is_shared()) colorIndex += offset;
    return color_for_block(binary_search_blocks(colorIndex), colorIndex);
}

// variable_weight_size_colors.h:415-551 (incremental block walk; start is absolute)
template <typename B>
void LeafView::copy_colors(B& builder, u64 start, u64 count) const {
    u32 m = (u32)(start / kColorsPerMacro);
    u32 block = binary_search_blocks(start);
    u64 macroLast = macro_last_block(m);
    u64 macroWeight = d->macro[2 * m + 1];
    auto load = [&](u32 bi, u32& hdr, u32& col, u32& bpw) {
        const u64 b = d->blocks[bi];
        hdr = (u32)b;
        col = (u32)(b >> 32);
        bpw = hdr_bpw(hdr);
    };
    u32 hdr, col, bpw;
    load(block, hdr, col, bpw);
    for (u64 ci = start; ci < start + count; ++ci) {
        if ((u32)(ci / kColorsPerMacro) != m) {
            m++;
            block++;
            load(block, hdr, col, bpw);
            macroLast = macro_last_block(m);
            macroWeight = d->macro[2 * m + 1];
        }
        const u32 local = (u32)(ci % kColorsPerMacro);
        if (block != macroLast && hdr_color_index((u32)d->blocks[block + 1]) <= local) {
            block++;
            load(block, hdr, col, bpw);
        }
        CompressedColor c;
        c.bitsPerWeight = (u8)bpw;
        c.colorBits = col;
        if (bpw != 0) {
            const u32 localWeight = (local - hdr_color_index(hdr)) * bpw;
            c.weight = extract_bits(bpw, d->weights, macroWeight + hdr_weight_offset(hdr) + localWeight);
        }
        builder.add(c);
    }
}
template void LeafView::copy_colors<ColorLeafBuilder>(ColorLeafBuilder&, u64, u64) const;

// variable_weight_size_colors.h:560-589: MSB-first packing into 32-bit words
void ColorLeafBuilder::add_weight(u8 w, u8 bpw) {
    const u32 target = (targetBitPos + 32 - bpw) % 64;
    const u64 ww = (u64)w << target;
    if (target > (u32)(32 - bpw)) {
        weights.back() |= (u32)(ww >> 32);
        targetBitPos -= bpw;
    }
    if (target < 32) {
        weights.push_back(0u);
        weights.back() |= (u32)(ww & 0xFFFFFFFFu);
        targetBitPos = target;
    }
    numWeights += bpw;
}
// variable_weight_size_colors.h:590-619
void ColorLeafBuilder::add(const CompressedColor& c) {
    if (colorIndex % kColorsPerMacro == 0) {
        lastMacroStartColor = colorIndex;
        lastMacroStartWeight = numWeights;
        macros.push_back({numWeights, (u32)blocks.size()});
    }
    if (colorIndex % kColorsPerMacro == 0 || lastColorBits != c.colorBits || lastBpw != c.bitsPerWeight) {
        lastColorBits = c.colorBits;
        lastBpw = c.bitsPerWeight;
        blocks.push_back({c.bitsPerWeight, c.colorBits, numWeights - lastMacroStartWeight, colorIndex - lastMacroStartColor});
    }
    if (c.bitsPerWeight > 0) add_weight(c.weight, c.bitsPerWeight);
    colorIndex++;
}
// variable_weight_size_colors.h:620-647
void ColorLeafBuilder::add_large_single_color(F3 color, u32 n) {
    CompressedColor c;
    c.set_single_color(color);
    u32 left = n;
    while (left > 0) {
        const u32 inMacro = kColorsPerMacro - (colorIndex % kColorsPerMacro);
        const u32 take = std::min(inMacro, left);
        left -= take;
        add(c);
        colorIndex += take - 1;
    }
}
// variable_weight_size_colors.h:649-689
void ColorLeafBuilder::build(ColorLeaf& leaf) const {
    leaf = ColorLeaf{};
    leaf.offset = ColorLeaf::kUnique;
    leaf.valid = true;
    leaf.blocks.resize(blocks.size());
    for (size_t i = 0; i < blocks.size(); ++i)
        leaf.blocks[i] = ((u64)blocks[i].colorBits << 32) | make_block_header(blocks[i].weightOffset, blocks[i].bpw, blocks[i].index);
    leaf.weights.resize(weights.size());
    for (size_t i = 0; i < weights.size(); ++i) leaf.weights[i] = __builtin_bswap32(weights[i]);
    leaf.macro.resize(2 * macros.size());
    for (size_t i = 0; i < macros.size(); ++i) {
        leaf.macro[2 * i] = macros[i].startBlock;
        leaf.macro[2 * i + 1] = macros[i].startWeight;
    }
}

// hash_dag_colors.h:37-50
LeafView HashDAGColors::get_leaf(u32 handle) const {
    if (handle & 0x80000000u) {
        const ColorLeaf& l = leaves[handle & 0x7FFFFFFFu];
        return {&l, l.offset};
    }
    return {&mainLeaf, offsets[handle]};
}

}  // namespace orc
