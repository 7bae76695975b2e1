// sliders.h -- the four sliding-piece attack backends, selectable at run time
// (the reference picks one at compile time: switches.h:102-125).
//
//   KS     Kogge-Stone occluded fills (MoveGeneratorBitboard.h:504-750).  No
//          tables.  Counting uses one fill per direction over ALL movers of
//          that direction at once: rays of different pieces in one direction
//          never overlap, so popcount(fill & targets) is the exact move count.
//   Plain  per-square magics, private 4096/512-entry windows, 2.25 MiB
//          (MoveGeneratorBitboard.h:810-827; magics found by a seeded search).
//   Fancy  fixed-shift fancy magics sharing one 97,264-entry table, 760 KiB
//          (MoveGeneratorBitboard.h:753-790; published constants).
//   Byte   the same magics indexing a u8 table of per-square unique-attack
//          slots, 146 KiB in total (MoveGeneratorBitboard.h:762-770); on the
//          GPU every workgroup stages it in LDS.
//
// Every backend returns attacks including the first blocker in each direction;
// callers mask own pieces.
#pragma once
#include "board.h"

namespace pf {

enum Backend : int { kKS = 0, kPlain = 1, kFancy = 2, kByte = 3 };

// Per-square magic entry: relevant-occupancy mask, multiplier, table window.
struct alignas(32) MagicEntry {
    u64 mask;
    u64 factor;
    u32 pos;   // window start in the shared (or private) table
    u32 off;   // Byte: first slot of this square's unique-attack list
    u64 pad;
};
static_assert(sizeof(MagicEntry) == 32, "MagicEntry is 32 bytes");

// All slider tables, laid out as one contiguous blob so that it can be copied
// to the device (or into LDS for the Byte backend) in one go.
struct SliderTables {
    MagicEntry plain_b[64], plain_r[64];  // plain: pos = sq*512 / sq*4096
    MagicEntry fancy_b[64], fancy_r[64];  // fancy + byte
    u64 plain_btab[64 * 512];
    u64 plain_rtab[64 * 4096];
    u64 fancy_tab[97264];
    // Byte-lookup blob: per-square byte indices, then the attack sets
    u8 byte_tab[97264];
    u8 byte_pad[16];  // keeps the u64 lists 16-byte aligned
    u64 byte_r[4900];
    u64 byte_b[1428];
};

struct KSBackend {
    static constexpr bool kSetwise = true;
    PF_HD u64 rook(int s, u64 occ) const { return rook_fill(bit(s), ~occ); }
    PF_HD u64 bishop(int s, u64 occ) const { return bishop_fill(bit(s), ~occ); }
    PF_HD u64 rook_multi(u64 set, u64 occ) const { return rook_fill(set, ~occ); }
    PF_HD u64 bishop_multi(u64 set, u64 occ) const { return bishop_fill(set, ~occ); }
};

struct PlainBackend {
    static constexpr bool kSetwise = false;
    const MagicEntry* eb;
    const MagicEntry* er;
    const u64* tb;
    const u64* tr;
    PF_HD u64 rook(int s, u64 occ) const {
        const MagicEntry& e = er[s];
        return tr[e.pos + (u32)(((occ & e.mask) * e.factor) >> 52)];
    }
    PF_HD u64 bishop(int s, u64 occ) const {
        const MagicEntry& e = eb[s];
        return tb[e.pos + (u32)(((occ & e.mask) * e.factor) >> 55)];
    }
    PF_HD u64 rook_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= rook(ctz(set), occ); set &= set - 1; }
        return a;
    }
    PF_HD u64 bishop_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= bishop(ctz(set), occ); set &= set - 1; }
        return a;
    }
};

struct FancyBackend {
    static constexpr bool kSetwise = false;
    const MagicEntry* eb;
    const MagicEntry* er;
    const u64* tab;
    PF_HD u64 rook(int s, u64 occ) const {
        const MagicEntry& e = er[s];
        return tab[e.pos + (u32)(((occ & e.mask) * e.factor) >> 52)];
    }
    PF_HD u64 bishop(int s, u64 occ) const {
        const MagicEntry& e = eb[s];
        return tab[e.pos + (u32)(((occ & e.mask) * e.factor) >> 55)];
    }
    PF_HD u64 rook_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= rook(ctz(set), occ); set &= set - 1; }
        return a;
    }
    PF_HD u64 bishop_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= bishop(ctz(set), occ); set &= set - 1; }
        return a;
    }
};

struct ByteBackend {
    static constexpr bool kSetwise = false;
    const MagicEntry* eb;
    const MagicEntry* er;
    const u8* bt;
    const u64* ur;
    const u64* ub;
    PF_HD u64 rook(int s, u64 occ) const {
        const MagicEntry& e = er[s];
        return ur[e.off + bt[e.pos + (u32)(((occ & e.mask) * e.factor) >> 52)]];
    }
    PF_HD u64 bishop(int s, u64 occ) const {
        const MagicEntry& e = eb[s];
        return ub[e.off + bt[e.pos + (u32)(((occ & e.mask) * e.factor) >> 55)]];
    }
    PF_HD u64 rook_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= rook(ctz(set), occ); set &= set - 1; }
        return a;
    }
    PF_HD u64 bishop_multi(u64 set, u64 occ) const {
        u64 a = 0;
        while (set) { a |= bishop(ctz(set), occ); set &= set - 1; }
        return a;
    }
};

}  // namespace pf
