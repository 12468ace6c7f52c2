/* oexpr.h -- oracle expression evaluator (test infrastructure only; see oexpr.c). */
#ifndef ORACLE_OEXPR_H
#define ORACLE_OEXPR_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oexpr oexpr;

/* Compile src against the symbol names[0..nnames-1]; NULL + message on error. */
oexpr* oexpr_compile(const char* src, const char* const* names, int nnames, char* err, int errlen);
/* Evaluate with vars[i] holding the value of names[i]. */
double oexpr_eval(const oexpr* e, const double* vars);
/* 1 if the expression reads symbol index var. */
int oexpr_uses(const oexpr* e, int var);
void oexpr_free(oexpr* e);

#ifdef __cplusplus
}
#endif
#endif
