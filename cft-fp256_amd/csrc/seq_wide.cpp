// The wide scheduler: a pure program's blocks as bundles of independent
// operations over W lanes (seq_wide.hpp).
#include "seq_wide.hpp"

#include <string.h>

#include <algorithm>
#include <bitset>
#include <map>
#include <vector>

#include "seq_prog.hpp"

namespace {

constexpr int kMaxSlots = 2048;
constexpr int kVars = 32 + kMaxSlots;      // registers 0..31, then scratch slots
using VarSet = std::bitset<kVars>;

int var_slot(uint32_t s) { return 32 + (int)s; }

// which of a, b, c an opcode reads (host/src/softfloat.c:822-848); an unread
// field creates no dependence
unsigned reads_of(uint32_t op)
{
    switch (op) {
    case 0: return 7u;                    // fma
    case 1: case 2: return 5u;            // add, sub: a, c
    case 3: return 3u;                    // mul
    case 4: case 5: case 26: case 27: return 1u;
    case 11: return 7u;                   // select
    default: return 3u;                   // binary (a, b); unassigned read a, b too
    }
}

struct Item {
    bool loop = false;
    uint32_t lo = 0, hi = 0;              // block [lo, hi); loop: REPEAT lo, ENDREP hi
    std::vector<Item> body;
};

void parse(const cft_program *p, uint32_t lo, uint32_t hi, std::vector<Item> &out)
{
    uint32_t start = lo, pc = lo;
    auto flush = [&](uint32_t end) {
        if (end > start) {
            Item b;
            b.lo = start;
            b.hi = end;
            out.push_back(b);
        }
    };
    while (pc < hi) {
        const SeqInsn &in = p->code[pc];
        if (SEQ_W_CTRL(in.w) && SEQ_W_OP(in.w) == SEQ_REPEAT) {
            flush(pc);
            Item L;
            L.loop = true;
            L.lo = pc;
            L.hi = in.aux;
            parse(p, pc + 1, in.aux, L.body);
            out.push_back(L);
            pc = in.aux + 1;
            start = pc;
            continue;
        }
        if (SEQ_W_CTRL(in.w) && SEQ_W_OP(in.w) == SEQ_HALT) {
            flush(pc);
            start = pc = hi;              // the rest is unreachable
            break;
        }
        pc++;
    }
    flush(pc);
}

// the program up to its first top-level HALT
uint32_t program_end(const cft_program *p)
{
    int depth = 0;
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const SeqInsn &in = p->code[pc];
        if (!SEQ_W_CTRL(in.w))
            continue;
        const uint32_t op = SEQ_W_OP(in.w);
        if (op == SEQ_REPEAT) depth++;
        else if (op == SEQ_ENDREP) depth--;
        else if (op == SEQ_HALT && depth == 0) return pc;
    }
    return p->n_insns;
}

void block_sets(const cft_program *p, uint32_t lo, uint32_t hi, VarSet &ue, VarSet &def)
{
    ue.reset();
    def.reset();
    for (uint32_t pc = lo; pc < hi; pc++) {
        const SeqInsn &in = p->code[pc];
        int rd_var = -1;
        if (!SEQ_W_CTRL(in.w)) {
            const unsigned rm = reads_of(SEQ_W_OP(in.w));
            const uint32_t s[3] = {SEQ_SRC_A(in.src), SEQ_SRC_B(in.src), SEQ_SRC_C(in.src)};
            const bool k[3] = {SEQ_W_KA(in.w) != 0, SEQ_W_KB(in.w) != 0, SEQ_W_KC(in.w) != 0};
            for (int i = 0; i < 3; i++)
                if ((rm >> i & 1u) && !k[i] && !def[s[i]])
                    ue.set(s[i]);
            rd_var = (int)SEQ_W_RD(in.w);
        } else if (SEQ_W_OP(in.w) == SEQ_LDL) {
            const int v = var_slot(in.imm & 0xFFFFFFu);
            if (!def[v])
                ue.set(v);
            rd_var = (int)SEQ_W_RD(in.w);
        } else if (SEQ_W_OP(in.w) == SEQ_STL) {
            const uint32_t ra = SEQ_SRC_A(in.src);
            if (!def[ra])
                ue.set(ra);
            def.set(var_slot(in.imm & 0xFFFFFFu));
        }
        if (rd_var >= 0)
            def.set(rd_var);
    }
}

// Backward liveness over a sequence of items; records each block's live-out.
VarSet live_in_seq(const cft_program *p, const std::vector<Item> &items, VarSet live_out,
                   std::map<uint32_t, VarSet> &block_out)
{
    VarSet L = live_out;
    for (auto it = items.rbegin(); it != items.rend(); ++it) {
        if (!it->loop) {
            block_out[it->lo] = L;
            VarSet ue, def;
            block_sets(p, it->lo, it->hi, ue, def);
            L = ue | (L & ~def);
        } else {
            // the body runs at least once; after it, the loop again or the exit
            VarSet body_out = L, body_in;
            for (int guard = 0; guard < 64; guard++) {
                body_in = live_in_seq(p, it->body, body_out, block_out);
                const VarSet next = L | body_in;
                if (next == body_out)
                    break;
                body_out = next;
            }
            L = body_in;
        }
    }
    return L;
}

// a source operand inside a block
struct Val {
    int kind = 0;          // 0 none, 1 constant, 2 node value, 3 entry value of a register
    uint32_t idx = 0;
    bool operator==(const Val &o) const { return kind == o.kind && idx == o.idx; }
    bool operator!=(const Val &o) const { return !(*this == o); }
};

struct Node {
    int kind = WK_ALU;
    uint32_t op = 0, rnd = 0;
    Val src[3];
    uint32_t mod[3] = {0, 0, 0};  // sign modifier on load: bit 0 clear, bit 1 flip
    uint32_t slot = 0;          // LOAD/STORE
    int home = -1;              // writes this home register directly (ALU, LOAD, MOVE)
    std::vector<int> raw, war;  // predecessors
    std::vector<int> users;     // RAW successors
    int bundle = -1;
    int prio = 0;
    int loc = -1;               // register its value lives in
};

struct BlockOut {
    std::vector<std::vector<int>> bundles;   // node ids
    std::vector<Node> nodes;
    uint32_t temps = 0;
};

// ABS and NEG (host/src/softfloat.c:954-961) change the sign bit alone and
// raise no flag: as a sign modifier on a reader's operand they give it the
// same bits.  A modifier is (clear, flip): x -> (clear ? |x| : x) ^ flip.
uint32_t sign_mod_of(uint32_t op)
{
    return op == 4 ? 1u : op == 5 ? 2u : 0u;
}

// outer(inner(x)) as one modifier
uint32_t compose_mod(uint32_t outer, uint32_t inner)
{
    if (outer & 1u)
        return 1u | (outer & 2u);
    return (inner & 1u) | ((inner ^ outer) & 2u);
}

// An ALU operand that is the result of an ABS or NEG in the block reads that
// instruction's own operand, with the sign change as a modifier on load: the
// ABS/NEG leaves the dependence chain, and is dropped when nothing else
// (a store, a register live after the block, a reader that is not an ALU
// operation) needs its result.  Nodes are in creation order, so a node's
// sources are final before its readers are rewritten.
void fuse_sign_ops(std::vector<Node> &N, std::map<uint32_t, Val> &stored,
                   std::map<uint32_t, int> &entry_load, std::vector<Val> &cur,
                   const VarSet &live_out, uint32_t nregs)
{
    // A register whose home is written in the block: a reader of its entry
    // value must come before that write (a WAR edge).  An ABS/NEG of such an
    // entry value stays, so no reader that depends on the write is made to
    // read the entry value (which would order it both before and after it).
    std::vector<bool> rewritten(32, false);
    for (uint32_t r = 0; r < nregs; r++)
        rewritten[r] = live_out[r] && !(cur[r] == Val{3, r});
    for (size_t i = 0; i < N.size(); i++) {
        if (N[i].kind != WK_ALU)
            continue;
        for (int k = 0; k < 3; k++) {
            const Val v = N[i].src[k];
            if (v.kind != 2)
                continue;
            const Node &q = N[v.idx];
            const uint32_t m = q.kind == WK_ALU ? sign_mod_of(q.op) : 0u;
            if (!m || (q.src[0].kind == 3 && rewritten[q.src[0].idx]))
                continue;
            N[i].src[k] = q.src[0];
            N[i].mod[k] = compose_mod(compose_mod(N[i].mod[k], m), q.mod[0]);
        }
    }
    std::vector<int> refs(N.size(), 0);
    for (const Node &n : N)
        for (int k = 0; k < 3; k++)
            if (n.src[k].kind == 2)
                refs[n.src[k].idx]++;
    for (auto &kv : stored)
        if (kv.second.kind == 2)
            refs[kv.second.idx]++;
    for (uint32_t r = 0; r < nregs; r++)
        if (live_out[r] && cur[r].kind == 2)
            refs[cur[r].idx]++;
    std::vector<int> remap(N.size(), -1);
    size_t kept = 0;
    for (size_t i = 0; i < N.size(); i++) {
        const bool dead = N[i].kind == WK_ALU && sign_mod_of(N[i].op) && refs[i] == 0;
        if (!dead)
            remap[i] = (int)kept++;
    }
    if (kept == N.size())
        return;
    std::vector<Node> M;
    M.reserve(kept);
    for (size_t i = 0; i < N.size(); i++)
        if (remap[i] >= 0)
            M.push_back(N[i]);
    auto fix = [&](Val &v) {
        if (v.kind == 2)
            v.idx = (uint32_t)remap[v.idx];
    };
    for (Node &n : M)
        for (int k = 0; k < 3; k++)
            fix(n.src[k]);
    for (auto &kv : stored)
        fix(kv.second);
    for (auto &kv : entry_load)
        kv.second = remap[kv.second];
    for (Val &v : cur)
        if (v.kind == 2 && remap[v.idx] < 0)
            v = Val{};                      // a dead value no live register holds
        else
            fix(v);
    N.swap(M);
}

int weight(const Node &n)
{
    if (n.kind != WK_ALU)
        return 1;
    return n.op <= 3 ? 4 : 1;
}

bool compile_block(const cft_program *p, uint32_t lo, uint32_t hi, const VarSet &live_out,
                   int W, int window, BlockOut &B)
{
    std::vector<Node> &N = B.nodes;
    N.clear();
    const uint32_t nregs = p->nregs;
    std::vector<Val> cur(32);
    for (uint32_t r = 0; r < 32; r++) {
        cur[r].kind = 3;
        cur[r].idx = r;
    }
    std::map<uint32_t, Val> stored;          // slot -> value stored in this block
    std::map<uint32_t, int> entry_load;      // slot -> LOAD node
    for (uint32_t pc = lo; pc < hi; pc++) {
        const SeqInsn &in = p->code[pc];
        if (!SEQ_W_CTRL(in.w)) {
            Node n;
            n.kind = WK_ALU;
            n.op = SEQ_W_OP(in.w);
            n.rnd = SEQ_W_RND(in.w);
            const unsigned rm = reads_of(n.op);
            const uint32_t s[3] = {SEQ_SRC_A(in.src), SEQ_SRC_B(in.src), SEQ_SRC_C(in.src)};
            const bool k[3] = {SEQ_W_KA(in.w) != 0, SEQ_W_KB(in.w) != 0, SEQ_W_KC(in.w) != 0};
            for (int i = 0; i < 3; i++) {
                if (k[i]) {
                    n.src[i].kind = 1;
                    n.src[i].idx = s[i];
                } else if (rm >> i & 1u) {
                    n.src[i] = cur[s[i]];
                }
            }
            N.push_back(n);
            cur[SEQ_W_RD(in.w)] = Val{2, (uint32_t)(N.size() - 1)};
            continue;
        }
        const uint32_t op = SEQ_W_OP(in.w);
        const uint32_t slot = in.imm & 0xFFFFFFu;
        if (op == SEQ_LDL) {
            auto st = stored.find(slot);
            if (st != stored.end()) {
                cur[SEQ_W_RD(in.w)] = st->second;
            } else {
                auto el = entry_load.find(slot);
                if (el == entry_load.end()) {
                    Node n;
                    n.kind = WK_LOAD;
                    n.slot = slot;
                    N.push_back(n);
                    entry_load[slot] = (int)N.size() - 1;
                    el = entry_load.find(slot);
                }
                cur[SEQ_W_RD(in.w)] = Val{2, (uint32_t)el->second};
            }
        } else if (op == SEQ_STL) {
            stored[slot] = cur[SEQ_SRC_A(in.src)];
        }
    }
    fuse_sign_ops(N, stored, entry_load, cur, live_out, nregs);
    // the stores and register values live after the block
    for (auto &kv : stored) {
        if (!live_out[var_slot(kv.first)])
            continue;
        auto el = entry_load.find(kv.first);
        if (el != entry_load.end() && kv.second == Val{2, (uint32_t)el->second})
            continue;                       // its own entry value: nothing to store
        Node n;
        n.kind = WK_STORE;
        n.slot = kv.first;
        n.src[0] = kv.second;
        N.push_back(n);
    }
    std::vector<int> home_writer(32, -1);
    for (uint32_t r = 0; r < nregs; r++) {
        if (!live_out[r] || cur[r] == Val{3, r})
            continue;
        const Val v = cur[r];
        if (v.kind == 2 && N[v.idx].home < 0 &&
            (N[v.idx].kind == WK_ALU || N[v.idx].kind == WK_LOAD)) {
            N[v.idx].home = (int)r;
            home_writer[r] = (int)v.idx;
        } else {
            Node n;
            n.kind = WK_MOVE;
            n.src[0] = v;
            n.home = (int)r;
            N.push_back(n);
            home_writer[r] = (int)N.size() - 1;
        }
    }
    // dependences
    for (size_t i = 0; i < N.size(); i++) {
        for (int k = 0; k < 3; k++) {
            const Val &v = N[i].src[k];
            if (v.kind == 2) {
                N[i].raw.push_back((int)v.idx);
                N[v.idx].users.push_back((int)i);
            } else if (v.kind == 3 && home_writer[v.idx] >= 0 &&
                       home_writer[v.idx] != (int)i) {
                N[home_writer[v.idx]].war.push_back((int)i);
            }
        }
        if (N[i].kind == WK_STORE) {
            auto el = entry_load.find(N[i].slot);
            if (el != entry_load.end())
                N[i].war.push_back(el->second);
        }
    }
    for (Node &n : N) {
        std::sort(n.raw.begin(), n.raw.end());
        n.raw.erase(std::unique(n.raw.begin(), n.raw.end()), n.raw.end());
        std::sort(n.war.begin(), n.war.end());
        n.war.erase(std::unique(n.war.begin(), n.war.end()), n.war.end());
    }
    // critical-path priorities (successors come later in creation order,
    // except a WAR edge whose writer was created earlier: handled by the
    // readiness test, not by priority)
    for (int i = (int)N.size() - 1; i >= 0; i--) {
        int best = 0;
        for (int u : N[i].users)
            best = std::max(best, N[u].prio);
        N[i].prio = best + weight(N[i]);
    }
    // list scheduling
    B.bundles.clear();
    size_t left = N.size(), oldest = 0;
    std::vector<int> cand;
    while (left) {
        while (oldest < N.size() && N[oldest].bundle >= 0)
            oldest++;
        const int b = (int)B.bundles.size();
        cand.clear();
        const size_t lim = std::min(N.size(), oldest + (size_t)window);
        for (size_t i = oldest; i < lim; i++) {
            if (N[i].bundle >= 0)
                continue;
            bool ok = true;
            for (int q : N[i].raw)
                if (N[q].bundle < 0 || N[q].bundle >= b) { ok = false; break; }
            if (ok)
                for (int q : N[i].war)
                    if (N[q].bundle < 0 || N[q].bundle >= b) { ok = false; break; }
            if (ok)
                cand.push_back((int)i);
        }
        if (cand.empty())
            return false;                    // cannot happen for a well-formed block
        std::stable_sort(cand.begin(), cand.end(),
                         [&](int x, int y) { return N[x].prio > N[y].prio; });
        // a scratch access sits at its slot's owner lane (slot mod W); any
        // other operation at any free lane
        std::vector<int> pos(W, -1);
        size_t taken = 0;
        for (int i : cand) {
            if (N[i].kind != WK_LOAD && N[i].kind != WK_STORE)
                continue;
            const int q = (int)(N[i].slot % (uint32_t)W);
            if (pos[q] < 0) {
                pos[q] = i;
                N[i].bundle = b;
                taken++;
            }
        }
        int free_pos = 0;
        for (int i : cand) {
            if (N[i].kind == WK_LOAD || N[i].kind == WK_STORE)
                continue;
            while (free_pos < W && pos[free_pos] >= 0)
                free_pos++;
            if (free_pos >= W)
                break;
            pos[free_pos] = i;
            N[i].bundle = b;
            taken++;
        }
        B.bundles.push_back(pos);
        left -= taken;
    }
    // temporaries: a value's register is free from the bundle after its last use
    std::vector<int> last(N.size(), -1);
    for (size_t i = 0; i < N.size(); i++) {
        int l = N[i].bundle;
        for (int u : N[i].users)
            l = std::max(l, N[u].bundle);
        last[i] = l;
    }
    std::vector<uint32_t> free_regs;
    std::vector<std::pair<int, uint32_t>> busy;       // (last use, register)
    uint32_t next = nregs;
    for (size_t b = 0; b < B.bundles.size(); b++) {
        for (size_t k = 0; k < busy.size();) {
            if (busy[k].first < (int)b) {
                free_regs.push_back(busy[k].second);
                busy[k] = busy.back();
                busy.pop_back();
            } else {
                k++;
            }
        }
        std::sort(free_regs.begin(), free_regs.end(), std::greater<uint32_t>());
        for (int i : B.bundles[b]) {
            if (i < 0)
                continue;
            Node &n = N[i];
            if (n.kind == WK_STORE)
                continue;
            if (n.home >= 0) {
                n.loc = n.home;
                continue;
            }
            uint32_t r;
            if (!free_regs.empty()) {
                r = free_regs.back();
                free_regs.pop_back();
            } else {
                r = next++;
            }
            n.loc = (int)r;
            busy.push_back({last[i], r});
        }
    }
    B.temps = next - nregs;
    return true;
}

uint32_t src_field(const std::vector<Node> &N, const Val &v, uint32_t &kbit)
{
    kbit = 0;
    switch (v.kind) {
    case 1: kbit = 1; return v.idx;
    case 2: return (uint32_t)N[v.idx].loc;
    case 3: return v.idx;
    default: return 0;
    }
}

struct Emitter {
    const cft_program *p;
    int W;
    int window;
    std::map<uint32_t, VarSet> block_out;
    WideProgram *out;
    uint32_t temps = 0;
    bool ok = true;

    void step_ctrl(uint32_t op, uint32_t y, uint32_t w)
    {
        WideRec r = {op | ((uint32_t)WK_CTRL << 8), y, 0, w};
        out->recs.push_back(r);
        for (int j = 1; j < W; j++)
            out->recs.push_back(WideRec{(uint32_t)WK_NOP << 8, 0, 0, 0});
        out->nsteps++;
    }

    void emit_block(uint32_t lo, uint32_t hi, uint32_t *alu)
    {
        BlockOut B;
        if (!compile_block(p, lo, hi, block_out[lo], W, window, B)) {
            ok = false;
            return;
        }
        temps = std::max(temps, B.temps);
        for (const std::vector<int> &bundle : B.bundles) {
            for (int j = 0; j < W; j++) {
                WideRec r = {(uint32_t)WK_NOP << 8, 0, 0, 0};
                if (bundle[j] >= 0) {
                    const Node &n = B.nodes[bundle[j]];
                    uint32_t ka, kb, kc;
                    const uint32_t a = src_field(B.nodes, n.src[0], ka);
                    const uint32_t b = src_field(B.nodes, n.src[1], kb);
                    const uint32_t c = src_field(B.nodes, n.src[2], kc);
                    switch (n.kind) {
                    case WK_ALU:
                        r.x = n.op | ((uint32_t)WK_ALU << 8) | (n.rnd << 11) | (ka << 14) |
                              (kb << 15) | (kc << 16) | (n.mod[0] << 17) | (n.mod[1] << 19) |
                              (n.mod[2] << 21);
                        r.y = a | (b << 16);
                        r.z = c | ((uint32_t)n.loc << 16);
                        if (alu) (*alu)++;
                        break;
                    case WK_LOAD:
                        r.x = (uint32_t)WK_LOAD << 8;
                        r.y = n.slot;
                        r.z = (uint32_t)n.loc << 16;
                        break;
                    case WK_STORE:
                        r.x = (uint32_t)WK_STORE << 8;
                        r.y = a;
                        r.z = n.slot << 16;
                        break;
                    default:
                        r.x = (uint32_t)WK_MOVE << 8;
                        r.y = a;
                        r.z = (uint32_t)n.loc << 16;
                        break;
                    }
                }
                out->recs.push_back(r);
            }
            out->nsteps++;
        }
    }

    void emit_seq(const std::vector<Item> &items, uint32_t *alu)
    {
        for (const Item &it : items) {
            if (!ok)
                return;
            if (!it.loop) {
                emit_block(it.lo, it.hi, alu);
                continue;
            }
            const uint32_t rep = out->nsteps;
            const bool top = it.lo == p->top_repeat;
            step_ctrl(SEQ_REPEAT, 0, p->code[it.lo].imm);
            if (top)
                out->top_step = rep;
            uint32_t body_alu = 0;
            emit_seq(it.body, top ? &body_alu : alu);
            if (top) {
                out->body_steps = out->nsteps - rep - 1;
                out->body_alu = body_alu;
            } else if (alu) {
                *alu += body_alu;
            }
            const uint32_t endrep = out->nsteps;
            step_ctrl(SEQ_ENDREP, rep + 1, 0);
            out->recs[(size_t)rep * W].y = endrep;
        }
    }
};

} // namespace

bool seq_wide_eligible(const cft_program *p)
{
    if (p->n_in > kMaxSlots || p->depth > (uint32_t)kMaxSlots)
        return false;
    const uint32_t end = program_end(p);
    for (uint32_t pc = 0; pc < end; pc++) {
        const SeqInsn &in = p->code[pc];
        if (!SEQ_W_CTRL(in.w))
            continue;
        switch (SEQ_W_OP(in.w)) {
        case SEQ_REPEAT: case SEQ_ENDREP: case SEQ_STL: case SEQ_LDL:
            break;
        default:
            return false;
        }
    }
    return true;
}

const int kWideWindows[] = {1 << 20, 4096, 2048, 1536, 1024, 768, 512, 384, 256,
                           128, 64, 32, 16};
const int kWideWindowCount = (int)(sizeof kWideWindows / sizeof kWideWindows[0]);

bool seq_wide_compile_at(const cft_program *p, int width, int window, WideProgram *out)
{
    std::vector<Item> items;
    parse(p, 0, program_end(p), items);
    WideProgram wp;
    wp.width = width;
    wp.nregs = p->nregs;
    Emitter E;
    E.p = p;
    E.W = width;
    E.window = std::max(window, width);
    E.out = &wp;
    VarSet exit_live;
    for (uint32_t s = 0; s < p->n_out && s < (uint32_t)kMaxSlots; s++)
        exit_live.set(var_slot(s));
    live_in_seq(p, items, exit_live, E.block_out);
    E.emit_seq(items, nullptr);
    if (!E.ok)
        return false;
    wp.nvregs = p->nregs + E.temps;
    *out = std::move(wp);
    return true;
}

bool seq_wide_compile(const cft_program *p, int width, uint32_t max_vregs, WideProgram *out)
{
    for (int i = 0; i < kWideWindowCount; i++) {
        WideProgram wp;
        if (!seq_wide_compile_at(p, width, kWideWindows[i], &wp))
            return false;
        if (wp.nvregs <= max_vregs) {
            *out = std::move(wp);
            return true;
        }
    }
    return false;
}
