-- A set-up file exercising the interpreted Lua subset: operator precedence
-- and associativity, modulo, concatenation, local and global functions,
-- recursion, closures, numeric for / while / repeat loops, tables,
-- short-circuit and/or, if / elseif / else and the math library.
local k = 3
G = 2 ^ 3 ^ 2           -- right associative: 2^9
N = -2 ^ 2              -- unary minus binds looser than ^
M = 7 % 3 + (-7) % 3    -- Lua's floored modulo: 1 + 2
S = "a" .. 1 .. "b"
local function fact(n)
    if n <= 1 then return 1 else return n * fact(n - 1) end
end
F = fact(5)
local acc = 0
for i = 1, 10, 3 do acc = acc + i end
for i = 10, 1, -4 do acc = acc + i end
local w = 0
while w < 5 do w = w + 2 end
repeat w = w + 1 until w > 7
function counter()
    local c = 0
    return function() c = c + 1; return c end
end
local nxt = counter()
nxt()
C2 = nxt()
T = {1, 2, x = 5, ["y"] = k}
local a, b = 1, 2
a, b = b, a

function initialise(x, y, z, xs, ys, zs)
    local r2 = (x - xs)^2 + (y - ys)^2 + (z - zs)^2
    local den
    if r2 < 1e36 then
        den = 1e-22
    elseif r2 < 4e36 then
        den = 2e-22
    else
        den = 3e-22
    end
    local pre = den * 8.314462e7 * 300
    local hii = (r2 < 1e35) and 1 or 0.25
    local checks = G + N + M + F + acc + w + C2 + #T + T.x + T.y + a * 10 + b
    if S == "a1b" then checks = checks + 1000 end
    return den, pre, hii, math.sqrt(r2) * 1e-18, math.log(r2 + 1, 10) + math.max(1, 2, 3) - math.min(4, 5),
        checks + math.floor(-2.5) + math.abs(-1) + math.pow(2, 0.5) * math.cos(math.pi) + math.atan2(1, 1),
        0, 0, 0
end
