// params.cpp -- key=value parameters of the command line and of a parameter
// file, with the lookup rules of the reference's getpar (src/cstd.c:936-1432):
//
//  * the table holds the key=value words of the file named by the last
//    non-empty par= word, then every key=value word of the command line, in
//    that order; words without '=' are ignored (so "freqs=0.25 ,0.75" gives
//    one frequency, as in the reference);
//  * file words are split at white space outside single or double quotes
//    (the quotes stay in the value), bytes are masked to 7 bits;
//  * a lookup finds the last occurrence of the key;
//  * numbers are comma-separated lists converted item by item (strtol / strtod
//    of the text from the item's start); the count of items read is returned,
//    so an empty value reads as absent; countparval counts commas + 1;
//  * a string lookup returns the whole value, commas included.
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "emmi_cli.h"

int iproc = 0, nproc = 1;

namespace {
struct Entry {
  std::string key;
  char* val;  // NUL-terminated, owned by g_store
};

std::vector<Entry> g_table;
std::vector<char*> g_store;
std::string g_prog = "emmi_main";
int g_argc = 0;
char** g_argv = nullptr;
bool g_tabled = false;

char* keep(const std::string& s) {
  char* p = static_cast<char*>(malloc(s.size() + 1));
  memcpy(p, s.c_str(), s.size() + 1);
  g_store.push_back(p);
  return p;
}

void add_word(const std::string& w) {
  const size_t eq = w.find('=');
  if (eq == std::string::npos) return;
  g_table.push_back({w.substr(0, eq), keep(w.substr(eq + 1))});
}

void read_par_file(const char* name) {
  FILE* fp = fopen(name, "r");
  if (!fp) emmi_cli_err("cannot open parameter file par=%s", name);
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
  fclose(fp);
  std::string word;
  bool in_word = false, quoted = false;
  for (char raw : text) {
    const char c = static_cast<char>(raw & 0x7f);
    if (!in_word) {
      if (isgraph(static_cast<unsigned char>(c))) {
        in_word = true;
        word.assign(1, c);
      }
    } else if (!quoted && isspace(static_cast<unsigned char>(c))) {
      add_word(word);
      in_word = false;
    } else {
      word.push_back(c);
    }
    if (c == '\'' || c == '"') quoted = !quoted;
  }
  if (in_word) add_word(word);
}

void tabulate() {
  if (g_tabled) return;
  g_tabled = true;
  g_table.clear();
  const char* par = nullptr;
  for (int i = g_argc - 1; i > 0; --i)
    if (!strncmp(g_argv[i], "par=", 4) && g_argv[i][4]) {
      par = g_argv[i] + 4;
      break;
    }
  if (par) read_par_file(par);
  for (int i = 1; i < g_argc; ++i) add_word(g_argv[i]);
}

const char* lookup(const char* name) {
  if (g_argc <= 1) return nullptr;
  tabulate();
  for (size_t i = g_table.size(); i-- > 0;)
    if (g_table[i].key == name) return g_table[i].val;
  return nullptr;
}

// calls f(item_start) for each comma-separated item; returns the count
template <class F>
int each_item(const char* v, F f) {
  int n = 0;
  for (const char* p = v; *p; ++n) {
    f(p, n);
    while (*p && *p != ',') ++p;
    if (*p == ',') ++p;
  }
  return n;
}
}  // namespace

extern "C" {

void initargs(int argc, char** argv) {
  g_argc = argc;
  g_argv = argv;
  g_tabled = false;
  if (argc > 0 && argv[0]) g_prog = argv[0];
}

int getparint(char* name, int* ptr) {
  const char* v = lookup(name);
  if (!v) return 0;
  return each_item(v, [&](const char* p, int k) { ptr[k] = static_cast<int>(strtol(p, nullptr, 10)); });
}

int getparfloat(char* name, float* ptr) {
  const char* v = lookup(name);
  if (!v) return 0;
  return each_item(v, [&](const char* p, int k) { ptr[k] = static_cast<float>(strtod(p, nullptr)); });
}

int getpardouble(char* name, double* ptr) {
  const char* v = lookup(name);
  if (!v) return 0;
  return each_item(v, [&](const char* p, int k) { ptr[k] = strtod(p, nullptr); });
}

int getparstring(char* name, char** ptr) {
  const char* v = lookup(name);
  if (!v) return 0;
  *ptr = const_cast<char*>(v);
  return 1;
}

int getparstringarray(char* name, char** ptr) {
  const char* v = lookup(name);
  if (!v) return 0;
  return each_item(v, [&](const char* p, int k) {
    const char* e = p;
    while (*e && *e != ',') ++e;
    ptr[k] = keep(std::string(p, e));
  });
}

int countparval(char* name) {
  const char* v = lookup(name);
  if (!v) return 0;
  int n = 1;
  for (const char* p = v; *p; ++p) n += (*p == ',');
  return n;
}

void emmi_cli_err(const char* fmt, ...) {
  fflush(stdout);
  fprintf(stderr, "\n%s: ", g_prog.c_str());
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  exit(EXIT_FAILURE);
}
}
