/* ORACLE — test infrastructure only. Nothing in the product links this.
 *
 * The reference's own Lua front end in a library: a Lua 5.2 state (the reference's
 * vendored lua-5.2.1, compiled where it lies) with the reference's `Mara` module
 * (Mara/mara.cpp, luaopen_Mara, and every Mara source it links) and its `buffer` module
 * (ctf-core/lua-buffer.c) preloaded. Tests drive it through plain C calls: run a chunk of
 * Lua, hand it double arrays as buffer userdata (the type Mara.set_primitive and
 * Mara.advance take), and read numbers back from a global table. This is how a Lua driver
 * script calls Mara (modules/MaraSim.lua:39-52), so what comes back is the reference's
 * answer through its own interface.
 *
 * Built by `make ref` into _ref/libmaralua.so with the reference sources it needs.
 */
#include <string.h>

#include "lua.h"
#include "lualib.h"
#include "lauxlib.h"

int luaopen_Mara(lua_State *L);
int luaopen_buffer(lua_State *L);

/* A new state with the standard libraries and the Mara and buffer modules loaded as
 * globals (require 'Mara' / require 'buffer' also work). NULL when Lua cannot allocate. */
void *ref_lua_new(void)
{
  lua_State *L = luaL_newstate();
  if (!L) return NULL;
  luaL_openlibs(L);
  luaL_requiref(L, "buffer", luaopen_buffer, 1);
  lua_pop(L, 1);
  luaL_requiref(L, "Mara", luaopen_Mara, 1);
  lua_pop(L, 1);
  return L;
}

/* Loads one more module into the state as global `name`: openf is its luaopen_ function
 * (the Lua binding of another library, such as the one INTEGRATION.md shows). */
int ref_lua_require(void *h, const char *name, void *openf)
{
  lua_State *L = (lua_State *)h;
  if (!openf) return 1;
  luaL_requiref(L, name, (lua_CFunction)openf, 1);
  lua_pop(L, 1);
  return 0;
}

void ref_lua_close(void *h)
{
  if (h) lua_close((lua_State *)h);
}

/* Runs a chunk. Returns 0, or 1 with Lua's error message copied into err. */
int ref_lua_run(void *h, const char *code, char *err, int nerr)
{
  lua_State *L = (lua_State *)h;
  int rc = luaL_loadstring(L, code);
  if (rc == LUA_OK) rc = lua_pcall(L, 0, 0, 0);
  if (rc != LUA_OK) {
    const char *m = lua_tostring(L, -1);
    if (err && nerr > 0) {
      strncpy(err, m ? m : "(no message)", (size_t)nerr - 1);
      err[nerr - 1] = '\0';
    }
    lua_pop(L, 1);
    return 1;
  }
  return 0;
}

/* Global `name` = a buffer userdata holding n doubles copied from x (luaL_setmetatable
 * 'buffer', as buffer.new_buffer makes them). */
int ref_lua_set_buffer(void *h, const char *name, const double *x, size_t n)
{
  lua_State *L = (lua_State *)h;
  void *p = lua_newuserdata(L, n * sizeof(double));
  if (!p) return 1;
  memcpy(p, x, n * sizeof(double));
  luaL_setmetatable(L, "buffer");
  lua_setglobal(L, name);
  return 0;
}

/* Copies the doubles of the buffer userdata in global `name` into x (at most n). Returns
 * the number copied, or -1 when the global is not a userdata. */
long ref_lua_get_buffer(void *h, const char *name, double *x, size_t n)
{
  lua_State *L = (lua_State *)h;
  lua_getglobal(L, name);
  if (!lua_isuserdata(L, -1)) {
    lua_pop(L, 1);
    return -1;
  }
  size_t m = lua_rawlen(L, -1) / sizeof(double);
  if (m > n) m = n;
  memcpy(x, lua_touserdata(L, -1), m * sizeof(double));
  lua_pop(L, 1);
  return (long)m;
}

/* Copies the numbers t[1..n] of the global table `name` into x. Returns the table's length,
 * or -1 when the global is not a table. */
long ref_lua_get_numbers(void *h, const char *name, double *x, size_t n)
{
  lua_State *L = (lua_State *)h;
  lua_getglobal(L, name);
  if (!lua_istable(L, -1)) {
    lua_pop(L, 1);
    return -1;
  }
  size_t len = lua_rawlen(L, -1);
  for (size_t i = 0; i < n && i < len; ++i) {
    lua_rawgeti(L, -1, (int)i + 1);
    x[i] = lua_tonumber(L, -1);
    lua_pop(L, 1);
  }
  lua_pop(L, 1);
  return (long)len;
}
