// Reader for the UHStencil DSL (the .dsl files of UH-GPU/<case>/): one
// `stencil` block with its array declaration, iteration(T) and one kernel
// whose `expr` is a sum of coefficient * access terms.  It produces the
// offsets and coefficients in term order, which is all the sweep kernels
// need; it does not build an AST or an IR.
//
// Grammar accepted (a subset of UHStencil/include/Parser/Parser.h:66-409,
// tokens as UHStencil/include/Parser/Lexer.h:31-59,155-215: identifiers,
// numbers with '.' marking a float, '#' comments to end of line):
//   stencil NAME ( (double|float) ARR [N]... ) {
//       iteration ( N )  |  operation ( KNAME )  |  mpiTile(...)  |  mpiHalo(...)
//       kernel KNAME { tile(...) | swCacheAt(...) | domain([lo,hi]...) | expr { SUM } }
//   }
//   SUM  := ['-'] TERM (('+'|'-') TERM)*
//   TERM := NUM '*' ACC | ACC '*' NUM | ACC
//   ACC  := ARR ('[' IDENT (('+'|'-') INT)? ']')+      (positional offsets)
// A '-' between terms negates the next coefficient, which is exact, so the
// left-to-right sum is unchanged (Parser.h:437-468: '+'/'-' at one
// precedence, '*' above them, left associative).
#include "uhs.h"
#include "uhs_host.h"
#include "uhs_dsl.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Tok {
    enum Kind { END, ID, NUM, SYM } kind;
    std::string text;
    double num;
    bool is_int;
    int line;
};

struct Lexer {
    const char *p;
    int line = 1;
    explicit Lexer(const char *s) : p(s) {}
    Tok next()
    {
        for (;;) {
            while (*p && isspace((unsigned char)*p)) {
                if (*p == '\n') line++;
                p++;
            }
            if (*p == '#') {
                while (*p && *p != '\n') p++;
                continue;
            }
            break;
        }
        Tok t{Tok::END, "", 0.0, false, line};
        if (!*p) return t;
        if (isalpha((unsigned char)*p) || *p == '_') {
            const char *s = p;
            while (isalnum((unsigned char)*p) || *p == '_') p++;
            t.kind = Tok::ID;
            t.text.assign(s, p);
            return t;
        }
        if (isdigit((unsigned char)*p) || (*p == '.' && isdigit((unsigned char)p[1]))) {
            const char *s = p;
            bool dot = false;
            while (isdigit((unsigned char)*p) || *p == '.') dot |= (*p++ == '.');
            t.kind = Tok::NUM;
            t.text.assign(s, p);
            t.num = strtod(t.text.c_str(), nullptr);
            t.is_int = !dot;
            return t;
        }
        t.kind = Tok::SYM;
        t.text.assign(p, p + 1);
        p++;
        return t;
    }
};

struct Parser {
    Lexer lex;
    Tok cur;
    std::string err;
    explicit Parser(const char *s) : lex(s) { cur = lex.next(); }
    void adv() { cur = lex.next(); }
    bool is(const char *sym) const { return cur.kind == Tok::SYM && cur.text == sym; }
    bool isid(const char *id) const { return cur.kind == Tok::ID && cur.text == id; }
    bool fail(const char *what)
    {
        if (err.empty()) {
            char b[256];
            snprintf(b, sizeof b, "line %d: expected %s, found '%s'", cur.line, what,
                     cur.kind == Tok::END ? "end of text" : cur.text.c_str());
            err = b;
        }
        return false;
    }
    bool expect(const char *sym)
    {
        if (!is(sym)) return fail(sym);
        adv();
        return true;
    }
    bool integer(long long *v)
    {
        if (cur.kind != Tok::NUM || !cur.is_int) return fail("an integer");
        *v = (long long)cur.num;
        adv();
        return true;
    }
    // '(' INT (',' INT)* ')' -- values ignored (tile, mpiTile, swCacheAt ...)
    bool int_list()
    {
        if (!expect("(")) return false;
        for (;;) {
            long long v;
            if (is("[")) { // mpiHalo([a,b][c,d]...) style: adjacent bracket pairs
                while (is("[")) {
                    adv();
                    if (!integer(&v) || !expect(",") || !integer(&v) || !expect("]")) return false;
                }
            } else if (!integer(&v)) {
                return false;
            }
            if (is(",")) { adv(); continue; }
            return expect(")");
        }
    }
};

} // namespace

static bool parse_access(Parser &ps, const std::string &arr, int dims, int *off)
{
    if (ps.cur.kind != Tok::ID || ps.cur.text != arr) return ps.fail(("array '" + arr + "'").c_str());
    ps.adv();
    int k = 0;
    while (ps.is("[")) {
        ps.adv();
        if (ps.cur.kind != Tok::ID) return ps.fail("an index name");
        ps.adv();
        int o = 0;
        if (ps.is("+") || ps.is("-")) {
            const int sign = ps.is("-") ? -1 : 1;
            ps.adv();
            long long v;
            if (!ps.integer(&v)) return false;
            o = sign * (int)v;
        }
        if (!ps.expect("]")) return false;
        if (k >= dims) return ps.fail("no more indices than the array has dimensions");
        off[k++] = o;
    }
    if (k != dims) return ps.fail("one index per array dimension");
    return true;
}

static bool parse_expr(Parser &ps, const std::string &arr, int dims, uhs_dsl_info *out)
{
    double sign = 1.0;
    if (ps.is("-")) {
        sign = -1.0;
        ps.adv();
    }
    for (;;) {
        if (out->npoints >= UHS_MAX_POINTS) return ps.fail("at most 343 terms");
        int *off = out->offsets + out->npoints * dims;
        double c = 1.0;
        if (ps.cur.kind == Tok::NUM) {
            c = ps.cur.num;
            ps.adv();
            if (!ps.expect("*")) return false;
            if (!parse_access(ps, arr, dims, off)) return false;
        } else {
            if (!parse_access(ps, arr, dims, off)) return false;
            if (ps.is("*")) {
                ps.adv();
                if (ps.cur.kind != Tok::NUM) return ps.fail("a coefficient");
                c = ps.cur.num;
                ps.adv();
            }
        }
        out->coeffs[out->npoints++] = sign * c;
        if (ps.is("+") || ps.is("-")) {
            sign = ps.is("-") ? -1.0 : 1.0;
            ps.adv();
            continue;
        }
        if (ps.is("}")) return true;
        return ps.fail("'+', '-' or the end of expr (only sums of coefficient*access terms are read)");
    }
}

static bool parse_kernel(Parser &ps, const std::string &arr, int dims, uhs_dsl_info *out,
                         bool *have_domain)
{
    ps.adv(); // 'kernel'
    if (ps.cur.kind != Tok::ID) return ps.fail("a kernel name");
    snprintf(out->kernel, sizeof out->kernel, "%s", ps.cur.text.c_str());
    ps.adv();
    if (!ps.expect("{")) return false;
    bool have_expr = false;
    while (!ps.is("}")) {
        if (ps.isid("tile") || ps.isid("swCacheAt")) {
            ps.adv();
            if (!ps.int_list()) return false;
        } else if (ps.isid("domain")) {
            ps.adv();
            if (!ps.expect("(")) return false;
            int k = 0;
            while (ps.is("[")) {
                ps.adv();
                long long lo, hi;
                if (!ps.integer(&lo) || !ps.expect(",") || !ps.integer(&hi) || !ps.expect("]"))
                    return false;
                if (k >= dims) return ps.fail("one domain range per dimension");
                out->domain_lo[k] = lo;
                out->domain_hi[k] = hi;
                k++;
            }
            if (k != dims) return ps.fail("one domain range per dimension");
            if (!ps.expect(")")) return false;
            *have_domain = true;
        } else if (ps.isid("expr")) {
            ps.adv();
            if (!ps.expect("{")) return false;
            if (!parse_expr(ps, arr, dims, out)) return false;
            if (!ps.expect("}")) return false;
            have_expr = true;
        } else {
            return ps.fail("tile, swCacheAt, domain or expr");
        }
    }
    ps.adv();
    if (!have_expr) return ps.fail("an expr block");
    return true;
}

int uhs_parse_dsl(const char *text, uhs_dsl_info *out)
{
    if (!text || !out) return uhs_set_error(UHS_E_INVAL, "null argument");
    memset(out, 0, sizeof *out);
    out->iteration = 1;
    Parser ps(text);
    if (!ps.isid("stencil")) return uhs_set_error(UHS_E_PARSE, ("line 1: expected 'stencil'"));
    ps.adv();
    if (ps.cur.kind != Tok::ID) { ps.fail("a stencil name"); return uhs_set_error(UHS_E_PARSE, ps.err.c_str()); }
    snprintf(out->name, sizeof out->name, "%s", ps.cur.text.c_str());
    ps.adv();
    std::string arr;
    int dims = 0, kernels = 0;
    bool have_domain = false;
    auto bail = [&]() { return uhs_set_error(UHS_E_PARSE, ps.err.c_str()); };
    if (!ps.expect("(")) return bail();
    if (ps.isid("double")) out->is_float = 0;
    else if (ps.isid("float")) out->is_float = 1;
    else { ps.fail("double or float"); return bail(); }
    ps.adv();
    if (ps.cur.kind != Tok::ID) { ps.fail("an array name"); return bail(); }
    arr = ps.cur.text;
    ps.adv();
    while (ps.is("[")) {
        ps.adv();
        long long n;
        if (!ps.integer(&n) || !ps.expect("]")) return bail();
        if (dims >= 3) { ps.fail("at most 3 dimensions"); return bail(); }
        out->grid[dims++] = n;
    }
    if (dims < 2) { ps.fail("a 2D or 3D array"); return bail(); }
    if (!ps.expect(")") || !ps.expect("{")) return bail();
    out->dims = dims;
    while (!ps.is("}")) {
        if (ps.cur.kind == Tok::END) { ps.fail("'}'"); return bail(); }
        if (ps.isid("iteration")) {
            ps.adv();
            long long t;
            if (!ps.expect("(") || !ps.integer(&t) || !ps.expect(")")) return bail();
            out->iteration = (int)t;
        } else if (ps.isid("operation")) {
            ps.adv();
            if (!ps.expect("(")) return bail();
            if (ps.cur.kind != Tok::ID) { ps.fail("a kernel name"); return bail(); }
            ps.adv();
            if (!ps.expect(")")) return bail();
        } else if (ps.isid("mpiTile") || ps.isid("mpiHalo")) {
            ps.adv();
            if (!ps.int_list()) return bail();
        } else if (ps.isid("kernel")) {
            if (kernels++ > 0)
                return uhs_set_error(UHS_E_UNSUPPORTED,
                                     "more than one kernel: read fused programs with uhs_parse_dsl_program");
            if (!parse_kernel(ps, arr, dims, out, &have_domain)) return bail();
        } else {
            ps.fail("iteration, operation, mpiTile, mpiHalo or kernel");
            return bail();
        }
    }
    if (!kernels) { ps.fail("a kernel"); return bail(); }
    // The reference requires iteration 1 or even (MLIRGen.cpp:317-321).
    if (out->iteration != 1 && out->iteration % 2)
        return uhs_set_error(UHS_E_PARSE, "iteration must be 1 or a multiple of 2");
    // Halo width = the domain's lower bound (UHAffineToParallel.cpp:1674-1680);
    // without a domain, the stencil radius.
    int radius = 0;
    for (int i = 0; i < out->npoints * dims; i++) radius = std::max(radius, std::abs(out->offsets[i]));
    for (int k = 0; k < dims; k++) {
        if (!have_domain) {
            out->domain_lo[k] = radius;
            out->domain_hi[k] = out->grid[k] - radius;
        }
        if (out->domain_lo[k] < radius || out->grid[k] - out->domain_hi[k] != out->domain_lo[k] ||
            out->domain_hi[k] <= out->domain_lo[k])
            return uhs_set_error(UHS_E_UNSUPPORTED,
                                 "the domain must be the interior with equal halos of at least "
                                 "the stencil radius");
        out->extents[k] = out->domain_hi[k] - out->domain_lo[k];
    }
    out->halo = (int)out->domain_lo[0];
    for (int k = 1; k < dims; k++)
        if (out->domain_lo[k] != out->halo)
            return uhs_set_error(UHS_E_UNSUPPORTED, "the halo must be equal on every axis");
    return UHS_OK;
}

int uhs_parse_dsl_file(const char *path, uhs_dsl_info *out)
{
    if (!path) return uhs_set_error(UHS_E_INVAL, "null path");
    FILE *f = fopen(path, "rb");
    if (!f) return uhs_set_error(UHS_E_IO, (std::string("cannot open ") + path).c_str());
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    return uhs_parse_dsl(text.c_str(), out);
}

// --------------------------------------------------------------------------
// Multi-kernel programs (uhs_parse_dsl_program): every kernel is read as a
// sum of coefficient * access terms whose accesses name the array or an
// earlier kernel (MLIRGen.cpp:356-432 resolves kernel results by name); the
// operation kernel (MLIRGen.cpp:187-188,259-264: its result is copied to the
// output) becomes the program's output, the kernels it reads become its
// producers (inlined, as StencilKernelFusion.cpp:66-180 does), and kernels
// it does not read are dropped, as they are dead in the reference too.
// --------------------------------------------------------------------------
namespace {

struct KTerm {
    std::string src;
    int off[3];
    double c;
};
struct KDef {
    std::string name;
    bool have_domain = false;
    long long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<KTerm> terms;
};

bool prog_access(Parser &ps, int dims, KTerm &t)
{
    if (ps.cur.kind != Tok::ID) return ps.fail("an array or kernel name");
    t.src = ps.cur.text;
    ps.adv();
    int k = 0;
    while (ps.is("[")) {
        ps.adv();
        if (ps.cur.kind != Tok::ID) return ps.fail("an index name");
        ps.adv();
        int o = 0;
        if (ps.is("+") || ps.is("-")) {
            const int sign = ps.is("-") ? -1 : 1;
            ps.adv();
            long long v;
            if (!ps.integer(&v)) return false;
            o = sign * (int)v;
        }
        if (!ps.expect("]")) return false;
        if (k >= dims) return ps.fail("no more indices than the array has dimensions");
        t.off[k++] = o;
    }
    if (k != dims) return ps.fail("one index per dimension");
    return true;
}

bool prog_kernel(Parser &ps, int dims, KDef &kd)
{
    ps.adv(); // 'kernel'
    if (ps.cur.kind != Tok::ID) return ps.fail("a kernel name");
    kd.name = ps.cur.text;
    ps.adv();
    if (!ps.expect("{")) return false;
    bool have_expr = false;
    while (!ps.is("}")) {
        if (ps.isid("tile") || ps.isid("swCacheAt")) {
            ps.adv();
            if (!ps.int_list()) return false;
        } else if (ps.isid("domain")) {
            ps.adv();
            if (!ps.expect("(")) return false;
            int k = 0;
            while (ps.is("[")) {
                ps.adv();
                long long lo, hi;
                if (!ps.integer(&lo) || !ps.expect(",") || !ps.integer(&hi) || !ps.expect("]")) return false;
                if (k >= dims) return ps.fail("one domain range per dimension");
                kd.lo[k] = lo;
                kd.hi[k] = hi;
                k++;
            }
            if (k != dims) return ps.fail("one domain range per dimension");
            if (!ps.expect(")")) return false;
            kd.have_domain = true;
        } else if (ps.isid("expr")) {
            ps.adv();
            if (!ps.expect("{")) return false;
            double sign = 1.0;
            if (ps.is("-")) {
                sign = -1.0;
                ps.adv();
            }
            for (;;) {
                KTerm t{};
                double c = 1.0;
                if (ps.cur.kind == Tok::NUM) {
                    c = ps.cur.num;
                    ps.adv();
                    if (!ps.expect("*") || !prog_access(ps, dims, t)) return false;
                } else {
                    if (!prog_access(ps, dims, t)) return false;
                    if (ps.is("*")) {
                        ps.adv();
                        if (ps.cur.kind != Tok::NUM) return ps.fail("a coefficient");
                        c = ps.cur.num;
                        ps.adv();
                    }
                }
                t.c = sign * c;
                kd.terms.push_back(t);
                if (ps.is("+") || ps.is("-")) {
                    sign = ps.is("-") ? -1.0 : 1.0;
                    ps.adv();
                    continue;
                }
                if (ps.is("}")) break;
                return ps.fail("'+', '-' or the end of expr (only sums of coefficient*access terms are read)");
            }
            ps.adv();
            have_expr = true;
        } else {
            return ps.fail("tile, swCacheAt, domain or expr");
        }
    }
    ps.adv();
    if (!have_expr) return ps.fail("an expr block");
    return true;
}

} // namespace

int uhs_parse_dsl_program(const char *text, uhs_dsl_program *out)
{
    if (!text || !out) return uhs_set_error(UHS_E_INVAL, "null argument");
    memset(out, 0, sizeof *out);
    out->iteration = 1;
    Parser ps(text);
    auto bail = [&]() { return uhs_set_error(UHS_E_PARSE, ps.err.c_str()); };
    if (!ps.isid("stencil")) return uhs_set_error(UHS_E_PARSE, "line 1: expected 'stencil'");
    ps.adv();
    if (ps.cur.kind != Tok::ID) { ps.fail("a stencil name"); return bail(); }
    snprintf(out->name, sizeof out->name, "%s", ps.cur.text.c_str());
    ps.adv();
    if (!ps.expect("(")) return bail();
    if (ps.isid("double")) out->is_float = 0;
    else if (ps.isid("float")) out->is_float = 1;
    else { ps.fail("double or float"); return bail(); }
    ps.adv();
    if (ps.cur.kind != Tok::ID) { ps.fail("an array name"); return bail(); }
    const std::string arr = ps.cur.text;
    ps.adv();
    int dims = 0;
    while (ps.is("[")) {
        ps.adv();
        long long n;
        if (!ps.integer(&n) || !ps.expect("]")) return bail();
        if (dims >= 3) { ps.fail("at most 3 dimensions"); return bail(); }
        out->grid[dims++] = n;
    }
    if (dims < 2) { ps.fail("a 2D or 3D array"); return bail(); }
    if (!ps.expect(")") || !ps.expect("{")) return bail();
    out->dims = dims;
    std::vector<KDef> kernels;
    std::string operation;
    while (!ps.is("}")) {
        if (ps.cur.kind == Tok::END) { ps.fail("'}'"); return bail(); }
        if (ps.isid("iteration")) {
            ps.adv();
            long long t;
            if (!ps.expect("(") || !ps.integer(&t) || !ps.expect(")")) return bail();
            out->iteration = (int)t;
        } else if (ps.isid("operation")) {
            ps.adv();
            if (!ps.expect("(")) return bail();
            if (ps.cur.kind != Tok::ID) { ps.fail("a kernel name"); return bail(); }
            operation = ps.cur.text;
            ps.adv();
            if (!ps.expect(")")) return bail();
        } else if (ps.isid("mpiTile") || ps.isid("mpiHalo")) {
            ps.adv();
            if (!ps.int_list()) return bail();
        } else if (ps.isid("kernel")) {
            KDef kd;
            if (!prog_kernel(ps, dims, kd)) return bail();
            for (const KDef &o : kernels)
                if (o.name == kd.name) return uhs_set_error(UHS_E_PARSE, ("kernel '" + kd.name + "' defined twice").c_str());
            kernels.push_back(std::move(kd));
        } else {
            ps.fail("iteration, operation, mpiTile, mpiHalo or kernel");
            return bail();
        }
    }
    if (kernels.empty()) { ps.fail("a kernel"); return bail(); }
    if (out->iteration != 1 && out->iteration % 2)
        return uhs_set_error(UHS_E_PARSE, "iteration must be 1 or a multiple of 2");
    if (operation.empty()) operation = kernels.back().name;
    int op = -1;
    for (int k = 0; k < (int)kernels.size(); k++)
        if (kernels[k].name == operation) op = k;
    if (op < 0) return uhs_set_error(UHS_E_PARSE, ("operation names no kernel: '" + operation + "'").c_str());
    snprintf(out->operation, sizeof out->operation, "%s", operation.c_str());
    const KDef &ok = kernels[op];
    if ((int)ok.terms.size() > UHS_MAX_PROGRAM_TERMS)
        return uhs_set_error(UHS_E_UNSUPPORTED, "more than 32 terms in the operation kernel");
    // Producers: every kernel the operation reads, directly or through other
    // kernels, in definition order (a kernel reads only the array and
    // kernels defined before it, so that order is a dependency order).
    auto index_of = [&](const std::string &name) {
        for (int k = 0; k < (int)kernels.size(); k++)
            if (kernels[k].name == name) return k;
        return -1;
    };
    std::vector<char> used(kernels.size(), 0);
    std::vector<int> stack{op};
    while (!stack.empty()) {
        const int k = stack.back();
        stack.pop_back();
        for (const KTerm &t : kernels[k].terms) {
            if (t.src == arr) continue;
            const int j = index_of(t.src);
            if (j < 0 || j >= k)
                return uhs_set_error(UHS_E_PARSE, ("kernel '" + kernels[k].name + "' reads '" + t.src +
                                                   "', which is neither the array nor an earlier kernel").c_str());
            if (!used[j]) {
                used[j] = 1;
                stack.push_back(j);
            }
        }
    }
    std::vector<int> prod, slot(kernels.size(), -1);
    for (int k = 0; k < op; k++)
        if (used[k]) {
            slot[k] = (int)prod.size();
            prod.push_back(k);
        }
    if ((int)prod.size() > UHS_MAX_PRODUCERS)
        return uhs_set_error(UHS_E_UNSUPPORTED, "the operation reads more than 4 kernels");
    // Reach per producer and of the program, per axis, below and above the
    // written cell (one-sided programs read only one side).
    int prlo[UHS_MAX_PRODUCERS][3], prhi[UHS_MAX_PRODUCERS][3];
    int rlo[3] = {0, 0, 0}, rhi[3] = {0, 0, 0};
    for (int i = 0; i < (int)prod.size(); i++) {
        const KDef &pk = kernels[prod[i]];
        if ((int)pk.terms.size() > UHS_MAX_PROGRAM_TERMS)
            return uhs_set_error(UHS_E_UNSUPPORTED, ("more than 32 terms in kernel '" + pk.name + "'").c_str());
        snprintf(out->producer_names[i], sizeof out->producer_names[i], "%s", pk.name.c_str());
        out->producer_npoints[i] = (int)pk.terms.size();
        for (int a = 0; a < dims; a++) { prlo[i][a] = INT32_MAX; prhi[i][a] = INT32_MIN; }
        for (int t = 0; t < (int)pk.terms.size(); t++) {
            const int s = pk.terms[t].src == arr ? -1 : slot[index_of(pk.terms[t].src)];
            out->producer_sources[i][t] = s;
            out->producer_coeffs[i][t] = pk.terms[t].c;
            for (int a = 0; a < dims; a++) {
                const int w = pk.terms[t].off[a];
                out->producer_offsets[i][t * dims + a] = w;
                prlo[i][a] = std::min(prlo[i][a], s < 0 ? w : w + prlo[s][a]);
                prhi[i][a] = std::max(prhi[i][a], s < 0 ? w : w + prhi[s][a]);
            }
        }
    }
    out->nproducers = (int)prod.size();
    out->npoints = (int)ok.terms.size();
    for (int t = 0; t < out->npoints; t++) {
        const KTerm &tm = ok.terms[t];
        const int src = tm.src == arr ? -1 : slot[index_of(tm.src)];
        out->sources[t] = src;
        out->coeffs[t] = tm.c;
        for (int a = 0; a < dims; a++) {
            out->offsets[t * dims + a] = tm.off[a];
            const int lo = src < 0 ? tm.off[a] : tm.off[a] + prlo[src][a];
            const int hi = src < 0 ? tm.off[a] : tm.off[a] + prhi[src][a];
            rlo[a] = std::max(rlo[a], -lo);
            rhi[a] = std::max(rhi[a], hi);
        }
    }
    // The written cells: the operation kernel's domain, or, when it has
    // none, the cells whose reads (below and above) stay inside the array.
    for (int a = 0; a < dims; a++) {
        out->domain_lo[a] = ok.have_domain ? ok.lo[a] : rlo[a];
        out->domain_hi[a] = ok.have_domain ? ok.hi[a] : out->grid[a] - rhi[a];
        if (out->domain_lo[a] < rlo[a] || out->domain_hi[a] > out->grid[a] - rhi[a] ||
            out->domain_hi[a] <= out->domain_lo[a])
            return uhs_set_error(UHS_E_UNSUPPORTED, "the operation's domain reads outside the array");
    }
    return UHS_OK;
}

int uhs_parse_dsl_program_file(const char *path, uhs_dsl_program *out)
{
    if (!path) return uhs_set_error(UHS_E_INVAL, "null path");
    FILE *f = fopen(path, "rb");
    if (!f) return uhs_set_error(UHS_E_IO, (std::string("cannot open ") + path).c_str());
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    return uhs_parse_dsl_program(text.c_str(), out);
}

int uhs_dsl_program_view(const uhs_dsl_program *dp, uhs_stencil *producers, uhs_program *pg)
{
    if (!dp || !pg || (dp->nproducers && !producers)) return uhs_set_error(UHS_E_INVAL, "null argument");
    memset(pg, 0, sizeof *pg);
    pg->dims = dp->dims;
    pg->nproducers = dp->nproducers;
    for (int i = 0; i < dp->nproducers; i++)
        producers[i] = uhs_stencil{dp->dims, dp->producer_npoints[i], dp->producer_offsets[i], dp->producer_coeffs[i]};
    pg->producers = producers;
    pg->output = uhs_stencil{dp->dims, dp->npoints, dp->offsets, dp->coeffs};
    pg->sources = dp->sources;
    bool chained = false;
    for (int i = 0; i < dp->nproducers; i++)
        for (int t = 0; t < dp->producer_npoints[i]; t++) chained |= dp->producer_sources[i][t] >= 0;
    pg->producer_sources = chained ? &dp->producer_sources[0][0] : nullptr;
    for (int a = 0; a < dp->dims; a++) {
        pg->lo[a] = dp->domain_lo[a];
        pg->hi[a] = dp->domain_hi[a];
    }
    return UHS_OK;
}
