// A minimal in-memory stand-in for the cosmosis datablock's C API, for testing the threepoint
// cosmosis module without cosmosis: typed values keyed by (section, name), the getters and
// setters the module calls, and helpers the tests use to fill and read blocks.  Test
// infrastructure only.
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {
struct Value {
    int type = 0;  // 1 int, 2 bool, 3 double, 4 string, 5 double array
    long long i = 0;
    double d = 0;
    std::string s;
    std::vector<double> arr;
    std::vector<int> ext;
};
struct Block {
    std::map<std::pair<std::string, std::string>, Value> v;
};
enum { OK = 0, NOT_FOUND = 5, WRONG_TYPE = 8, NDIM_MISMATCH = 15 };
Value *find(void *b, const char *sec, const char *name)
{
    auto &m = ((Block *)b)->v;
    auto it = m.find({sec, name});
    return it == m.end() ? nullptr : &it->second;
}
}  // namespace

extern "C" {
void *mock_block_new() { return new Block(); }
void mock_block_free(void *b) { delete (Block *)b; }
void mock_put_int(void *b, const char *s, const char *n, int v) { auto &x = ((Block *)b)->v[{s, n}]; x = Value(); x.type = 1; x.i = v; }
void mock_put_bool(void *b, const char *s, const char *n, int v) { auto &x = ((Block *)b)->v[{s, n}]; x = Value(); x.type = 2; x.i = v != 0; }
void mock_put_double(void *b, const char *s, const char *n, double v) { auto &x = ((Block *)b)->v[{s, n}]; x = Value(); x.type = 3; x.d = v; }
void mock_put_string(void *b, const char *s, const char *n, const char *v) { auto &x = ((Block *)b)->v[{s, n}]; x = Value(); x.type = 4; x.s = v; }
void mock_put_array(void *b, const char *s, const char *n, const double *v, int ndim, const int *ext)
{
    auto &x = ((Block *)b)->v[{s, n}];
    x = Value();
    x.type = 5;
    size_t total = 1;
    for (int i = 0; i < ndim; ++i) {
        x.ext.push_back(ext[i]);
        total *= (size_t)ext[i];
    }
    x.arr.assign(v, v + total);
}
int mock_get_array(void *b, const char *s, const char *n, double *out, int cap)
{
    Value *x = find(b, s, n);
    if (!x || x->type != 5) return -1;
    int m = (int)x->arr.size();
    for (int i = 0; i < m && i < cap; ++i) out[i] = x->arr[i];
    return m;
}

// the C API the module calls
int c_datablock_get_int(void *b, const char *s, const char *n, int *v)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 1) return WRONG_TYPE;
    *v = (int)x->i;
    return OK;
}
int c_datablock_get_bool(void *b, const char *s, const char *n, bool *v)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 2) return WRONG_TYPE;
    *v = x->i != 0;
    return OK;
}
int c_datablock_get_double(void *b, const char *s, const char *n, double *v)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type == 1) {
        *v = (double)x->i;
        return OK;
    }
    if (x->type != 3) return WRONG_TYPE;
    *v = x->d;
    return OK;
}
int c_datablock_get_string(void *b, const char *s, const char *n, char **v)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 4) return WRONG_TYPE;
    *v = (char *)std::malloc(x->s.size() + 1);
    std::memcpy(*v, x->s.c_str(), x->s.size() + 1);
    return OK;
}
int c_datablock_get_double_array_1d(void *b, const char *s, const char *n, double **v, int *size)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 5 || x->ext.size() != 1) return WRONG_TYPE;
    *size = (int)x->arr.size();
    *v = (double *)std::malloc(sizeof(double) * x->arr.size());
    std::memcpy(*v, x->arr.data(), sizeof(double) * x->arr.size());
    return OK;
}
int c_datablock_get_double_array_shape(void *b, const char *s, const char *n, int ndims, int *ext)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 5) return WRONG_TYPE;
    if ((int)x->ext.size() != ndims) return NDIM_MISMATCH;
    for (int i = 0; i < ndims; ++i) ext[i] = x->ext[i];
    return OK;
}
int c_datablock_get_double_array(void *b, const char *s, const char *n, double *v, int ndims, const int *ext)
{
    Value *x = find(b, s, n);
    if (!x) return NOT_FOUND;
    if (x->type != 5) return WRONG_TYPE;
    if ((int)x->ext.size() != ndims) return NDIM_MISMATCH;
    for (int i = 0; i < ndims; ++i)
        if (x->ext[i] != ext[i]) return NDIM_MISMATCH;
    std::memcpy(v, x->arr.data(), sizeof(double) * x->arr.size());
    return OK;
}
int c_datablock_put_double_array_1d(void *b, const char *s, const char *n, const double *v, int size)
{
    int e = size;
    mock_put_array(b, s, n, v, 1, &e);
    return OK;
}
}
