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

#include <cctype>
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
            if (is("[")) { // mpiHalo([a,b]...) style
                adv();
                if (!integer(&v) || !expect(",") || !integer(&v) || !expect("]")) return false;
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
                                     "more than one kernel (fused stencil programs) is not read");
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
